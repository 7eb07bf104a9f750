"""Sanitizer builds of the host code: the oracle driven through every entry point under
AddressSanitizer and UndefinedBehaviorSanitizer (CPU), and the C++ host-class test built with
the same sanitizers on its host code (run on the GPU). Device code is never built with a
sanitizer."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CPP = os.path.join(ROOT, "tests", "cpp")
SAN_ENV = dict(os.environ, UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    p = subprocess.run([os.path.join(ORACLE, "oracle_sanitize")], capture_output=True, text=True, timeout=600,
                       env=dict(SAN_ENV, OMP_NUM_THREADS="2"))
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-4000:]
    assert "no errors" in p.stdout


def test_host_test_builds_with_asan_ubsan():
    subprocess.run(["make", "-s", "-C", CPP, os.path.join(CPP, "test_convectiondiffusiondg_san")], check=True)
    assert os.path.exists(os.path.join(CPP, "test_convectiondiffusiondg_san"))


@pytest.mark.gpu
def test_host_test_under_asan_ubsan():
    """The host-class test with sanitized host code. Leak detection is off: the HIP runtime keeps
    its allocations until the process ends."""
    print("[gpu] C++ host-class test under ASan/UBSan (host code)", flush=True)
    binp = os.path.join(CPP, "test_convectiondiffusiondg_san")
    if not os.path.exists(binp):
        subprocess.run(["make", "-s", "-C", CPP, binp], check=True)
    env = dict(SAN_ENV, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0")
    p = subprocess.run([binp], capture_output=True, text=True, timeout=300, env=env)
    print(p.stdout[-2000:], p.stderr[-4000:], flush=True)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "PASSED" in p.stdout
