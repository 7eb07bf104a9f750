"""Adjoint gradient through the libtorch autograd Function; the checks run in a child
process (tests/adjoint_checks.py) that imports torch before the HIP library."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _run(which):
    p = subprocess.run([sys.executable, "-u", os.path.join(HERE, "adjoint_checks.py"), which],
                       capture_output=True, text=True, timeout=280)
    print(p.stdout, p.stderr[-3000:], flush=True)
    assert p.returncode == 0, p.stdout + p.stderr[-3000:]
    assert "ADJOINT CHECKS PASSED" in p.stdout


@pytest.mark.gpu
def test_autograd_gradient_vs_oracle():
    print("[gpu] autograd adjoint vs oracle 6^3 Q2", flush=True)
    _run("oracle")


@pytest.mark.gpu
def test_autograd_gradient_vs_finite_differences_32cube():
    print("[gpu] autograd adjoint vs central FD, 32^3 Q2, 20 cells", flush=True)
    _run("fd")
