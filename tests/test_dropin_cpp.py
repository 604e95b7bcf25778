"""Runs the C++ drop-in header test program (tests/cpp/test_dropin.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "build", "test_dropin")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return BIN


def test_dropin_header_exceptions_host_only():
    out = subprocess.run([_binary(), "--host-only"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "host-only checks OK" in out.stdout


@pytest.mark.gpu
def test_dropin_header_on_gpu():
    out = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "drop-in header checks OK" in out.stdout


ADAPTER = os.path.join(ROOT, "tests", "cpp", "build", "test_ddc_adapter")


def _adapter():
    if not os.path.exists(ADAPTER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    return ADAPTER


def test_ddc_adapter_host_only():
    """The DDC-typed overloads map each DDC domain kind onto its per-axis
    storage (compiled against the test double of DDC's names)."""
    out = subprocess.run([_adapter(), "--host-only"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "adapter host-only checks OK" in out.stdout


@pytest.mark.gpu
def test_ddc_adapter_on_gpu():
    out = subprocess.run([_adapter()], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ddc adapter checks OK" in out.stdout
