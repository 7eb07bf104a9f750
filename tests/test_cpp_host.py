"""Runs the C++ test written against the PDELab-shaped host classes
(dune-pdelab_amd/host/pdelab_amd.hh) -- the reference's own test style: a main()
that returns nonzero on failure."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_convectiondiffusiondg")


@pytest.mark.gpu
def test_cpp_convectiondiffusiondg():
    print("[gpu] C++ host-class test", flush=True)
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout


def test_cpp_host_header_compiles(tmp_path):
    """The host classes compile as plain C++17 against the C-ABI header (no GPU)."""
    src = tmp_path / "t.cc"
    src.write_text('#include "pdelab_amd.hh"\nint main(){return 0;}\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "dune-pdelab_amd", "host"),
                    str(src)], check=True)
