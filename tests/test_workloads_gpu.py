"""The benchmark workloads at full size on the GPU, checked against the CPU
oracle (OpenMP schedule, bit-identical to the serial one) and through
size-independent properties: round trip, mass conservation, and for the
float32 storage path its error against the float64 oracle.
"""
import numpy as np
import pytest
import torch

import oracle_bridge as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def gpu_transform(pw, field, level, lmin, name, forward=True):
    t = field.clone()
    (pw.hierarchize if forward else pw.dehierarchize)(t, level, lmin, level, name)
    torch.cuda.synchronize()
    return t


def oracle(a, level, lmin, name, forward=True):
    b = np.ascontiguousarray(a.copy())
    ob.transform(b, level, lmin, level, name, forward, ob.OMP_DIRECT)
    return b


def lattice(a, level):
    """The periodic lattice part of a grid stored with images."""
    return a[tuple(slice(0, 1 << l) for l in level)]


def test_hat_2d_1025(pw):
    from paliwa_amd import fields

    level, lmin = [10, 10], [0, 0]
    f = fields.gaussian_bump_2d(level, seed=0, device="cuda")
    host = f.cpu().numpy()
    fwd = gpu_transform(pw, f, level, lmin, "hat")
    want = oracle(host, level, lmin, "hat")
    assert np.array_equal(fwd.cpu().numpy(), want)
    back = gpu_transform(pw, fwd, level, lmin, "hat", forward=False)
    assert np.array_equal(back.cpu().numpy(), oracle(want, level, lmin, "hat", False))
    err = (back - f).abs().max().item() / f.abs().max().item()
    assert err <= 1e-14, err


@pytest.mark.parametrize("lmin", [[0, 0, 0], [2, 3, 1]])
def test_biorthogonal_3d_257(pw, lmin):
    from paliwa_amd import fields

    level = [8, 8, 8]
    f = fields.fourier_modes(level, modes=16, kmax=32, seed=1, device="cuda")
    host = f.cpu().numpy()
    fwd = gpu_transform(pw, f, level, lmin, "biorthogonal")
    want = oracle(host, level, lmin, "biorthogonal")
    assert np.array_equal(fwd.cpu().numpy(), want)
    back = gpu_transform(pw, fwd, level, lmin, "biorthogonal", forward=False)
    assert np.array_equal(back.cpu().numpy(), oracle(want, level, lmin, "biorthogonal", False))
    err = (back - f).abs().max().item() / f.abs().max().item()
    assert err <= 1e-14, err
    if lmin == [0, 0, 0]:
        # mass: the single coarse coefficient is the mean over the lattice
        mean = lattice(host, level).mean()
        coarse = fwd[0, 0, 0].item()
        assert abs(coarse - mean) <= 1e-14 * np.abs(host).max()


def fp32_errors(pw, level, lmin, name, f64):
    """float32 storage and arithmetic against the float64 oracle."""
    f32 = f64.to(torch.float32)
    fwd32 = gpu_transform(pw, f32, level, lmin, name)
    host64 = f64.cpu().numpy()
    fwd64 = oracle(host64, level, lmin, name)
    # bit-exact against the float32 oracle
    assert np.array_equal(fwd32.cpu().numpy(), oracle(f32.cpu().numpy(), level, lmin, name))
    d = fwd32.cpu().numpy().astype(np.float64) - fwd64
    max_rel = np.abs(d).max() / np.abs(fwd64).max()
    l2 = np.linalg.norm(d) / np.linalg.norm(fwd64)
    n_lattice = np.prod([1 << l for l in level])
    mass64 = lattice(host64, level).sum()
    mass32 = float(fwd32[(0,) * len(level)].item()) * n_lattice
    mass_rel = abs(mass32 - mass64) / np.abs(lattice(host64, level)).sum()
    back32 = gpu_transform(pw, fwd32, level, lmin, name, forward=False)
    rt = (back32.double() - f64).abs().max().item() / f64.abs().max().item()
    return max_rel, l2, mass_rel, rt


def test_float32_storage_against_float64(pw):
    from paliwa_amd import fields

    level, lmin = [8, 8, 8], [0, 0, 0]
    f64 = fields.fourier_modes(level, modes=16, kmax=32, seed=1, device="cuda")
    max_rel, l2, mass_rel, rt = fp32_errors(pw, level, lmin, "biorthogonal", f64)
    print(f"float32 vs float64 oracle: max-rel {max_rel:.3e}, L2 {l2:.3e}, "
          f"mass {mass_rel:.3e}, round trip {rt:.3e}")
    # float32 has a 2^-24 ~ 6e-8 unit roundoff; 8 levels x 3 axes x 2 passes
    assert max_rel < 1e-6
    assert l2 < 1e-6
    assert mass_rel < 1e-6
    assert rt < 1e-6


@pytest.mark.parametrize("name", ["hat", "biorthogonal"])
def test_anisotropic_6d(pw, name):
    from paliwa_amd import fields

    level, lmin = [5, 5, 4, 4, 3, 3], [0, 0, 0, 0, 0, 0]
    f = fields.cosine_modes_noise(level, seed=3, device="cuda")
    assert f.numel() == 25492401
    host = f.cpu().numpy()
    fwd = gpu_transform(pw, f, level, lmin, name)
    want = oracle(host, level, lmin, name)
    assert np.array_equal(fwd.cpu().numpy(), want)
    back = gpu_transform(pw, fwd, level, lmin, name, forward=False)
    assert np.array_equal(back.cpu().numpy(), oracle(want, level, lmin, name, False))
    # The round trip equals the reference arithmetic's own (bit-identical to
    # the oracle above); over 6 axes it accumulates to ~1.4e-14 of max|f|,
    # inside the reference test's 1e-13 (test/test_paliwa_transform.cpp:70-73).
    err = (back - f).abs().max().item() / f.abs().max().item()
    assert err <= 2e-14, err
