"""The C-ABI boundary without a GPU: the library loads, exports every entry point
include/pdelab_gpu.h declares, the ctypes struct layouts match the C compiler's, and
the calls that need no device fail cleanly (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdelab_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pdelab_gpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layout_matches_c(pkg, tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "pdelab_gpu.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(pdelab_config),'
                   ' offsetof(pdelab_config, gaussians), offsetof(pdelab_config, bctype),'
                   ' sizeof(pdelab_linear_solver_result), sizeof(pdelab_linear_problem_result),'
                   ' sizeof(pdelab_mixed_result), offsetof(pdelab_config, problem),'
                   ' offsetof(pdelab_config, j0), offsetof(pdelab_config, o0)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    C = pkg.config
    assert vals[0] == ctypes.sizeof(C.PdelabConfig)
    assert vals[1] == C.PdelabConfig.gaussians.offset
    assert vals[2] == C.PdelabConfig.bctype.offset
    assert vals[3] == ctypes.sizeof(C.LinearSolverResult)
    assert vals[4] == ctypes.sizeof(C.LinearProblemResult)
    assert vals[5] == ctypes.sizeof(C.MixedResult)
    assert vals[6] == C.PdelabConfig.problem.offset
    assert vals[7] == C.PdelabConfig.j0.offset
    assert vals[8] == C.PdelabConfig.o0.offset


def test_default_config_matches_reference_defaults(pkg):
    """ConvectionDiffusionDG defaults SIPG, weightsOn, alpha 1, intorderadd 0
    (R:dune/pdelab/localoperator/convectiondiffusiondg.hh:85-90)."""
    cfg = pkg.PdelabConfig()
    pkg.lib().pdelab_gpu_default_config(ctypes.byref(cfg))
    assert cfg.abi_version == pkg.config.ABI_VERSION == 2
    assert cfg.method == pkg.SIPG and cfg.weights == pkg.WEIGHTS_ON
    assert cfg.alpha == 1.0 and cfg.intorderadd == 0 and cfg.kappa0 == 1.0
    mine = pkg.make_config()
    assert (mine.method, mine.weights, mine.alpha) == (cfg.method, cfg.weights, cfg.alpha)


def test_create_rejects_bad_configs_before_touching_a_device(pkg):
    L = pkg.lib()
    h = ctypes.c_void_p()
    for kw in (dict(degree=0), dict(degree=4), dict(n=0), dict(method=7), dict(bctype=(5, 0, 0, 0, 0, 0)),
               dict(lx=-1.0), dict(kappa0=0.0)):
        cfg = pkg.make_config(**kw)
        st = L.pdelab_gpu_create(ctypes.byref(cfg), None, ctypes.byref(h))
        assert st in (1, 4), (kw, st)
        assert L.pdelab_gpu_last_error(None)
        assert not h.value
    cfg = pkg.make_config()
    cfg.abi_version = 99
    assert L.pdelab_gpu_create(ctypes.byref(cfg), None, ctypes.byref(h)) == 1
    cfg = pkg.make_config()
    cfg.kappa_per_cell = 3  # 0, 1 or 2 only
    assert L.pdelab_gpu_create(ctypes.byref(cfg), None, ctypes.byref(h)) == 1
    assert b"kappa_per_cell" in L.pdelab_gpu_last_error(None)


def test_kappa_layouts_checked_on_the_host(pkg):
    """kappa_per_cell = 2 (a diagonal tensor per cell) takes three values per cell; the binding
    checks the length before the library reads the array."""
    cfg = pkg.make_config(n=2, kappa_per_cell=2)
    assert cfg.kappa_per_cell == 2 and pkg.make_config(n=2, kappa_per_cell=True).kappa_per_cell == 1
    assert pkg.GridOperator._kappa_size(cfg) == 24
    with pytest.raises(ValueError, match="24"):
        pkg.GridOperator(cfg, [1.0] * 8)


def test_null_handles_are_invalid_arguments(pkg):
    L = pkg.lib()
    assert L.pdelab_gpu_jacobian_apply(None, None, None) == 1
    assert L.pdelab_gpu_residual(None, None, None) == 1
    assert L.pdelab_gpu_destroy(None) == 0
    assert L.pdelab_gpu_size(None) == 0


def test_package_fails_loudly_without_library(pkg, monkeypatch, tmp_path):
    monkeypatch.setattr(pkg, "_lib", None)
    monkeypatch.setattr(pkg, "LIBPATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="not built"):
        pkg.lib()
