"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host
code only): oracle/sanitize_driver.cpp drives every entry point of its C ABI
on small grids, both dtypes, every schedule, partial buffers and the
argument errors; any out-of-bounds access or undefined behaviour aborts it."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_address_and_ub_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    exe = os.path.join(ROOT, "oracle", "build", "sanitize_driver")
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "sanitizer checks OK" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
