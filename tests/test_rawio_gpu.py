"""Raw dumps from device memory (paliwa_dump_to_binary_file) read with numpy,
and loaded back, including a grid spanning several 16 MiB staging chunks."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rawio():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from paliwa_amd import rawio

    return rawio


@pytest.mark.parametrize("shape,dtype", [((17, 33), torch.float64), ((5, 9, 17), torch.float32),
                                         ((1025, 1025, 9), torch.float64)])
def test_dump_and_load(rawio, tmp_path, shape, dtype):
    g = torch.randn(shape, dtype=dtype, device="cuda")
    path = str(tmp_path / "g.raw")
    rawio.dump(g, path)
    want = g.cpu().numpy()
    assert np.array_equal(rawio.read_raw(path, shape, want.dtype), want)
    back = rawio.load(path, shape, dtype)
    assert torch.equal(back, g)


def test_errors(rawio, tmp_path):
    import paliwa_amd as pw

    g = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(OSError):
        rawio.dump(g, "/nonexistent-dir/g.raw")
    path = str(tmp_path / "g.raw")
    rawio.dump(g, path)
    with pytest.raises(pw.ShapeError):
        rawio.load(path, (9,))
