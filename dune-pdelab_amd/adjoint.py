"""torch.autograd bridge to the HIP solve (libtorch C++ Function in torch_ext/adjoint.cpp).

Build in-tree with ``build_extension()`` (links lib/libpdelab_gpu.so); ``load()`` imports
the built module. Import torch before loading the HIP library so both use torch's HIP
runtime (both carry the SONAME libamdhip64.so.7).
"""
import glob
import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
EXT_DIR = os.path.join(HERE, "torch_ext")
BUILD_DIR = os.path.join(EXT_DIR, "_build")
NAME = "pdelab_amd_adjoint"


def build_extension(verbose=False):
    import torch  # noqa: F401
    from torch.utils.cpp_extension import load
    os.makedirs(BUILD_DIR, exist_ok=True)
    lib = os.path.join(HERE, "lib")
    return load(name=NAME, sources=[os.path.join(EXT_DIR, "adjoint.cpp")],
                extra_include_paths=[os.path.join(HERE, "..", "include")],
                extra_ldflags=[f"-L{lib}", "-lpdelab_gpu", f"-Wl,-rpath,{lib}"],
                build_directory=BUILD_DIR, verbose=verbose, with_cuda=False)


def load():
    import torch  # noqa: F401  (first: fixes the HIP runtime of the process)
    if NAME in sys.modules:
        return sys.modules[NAME]
    cands = glob.glob(os.path.join(BUILD_DIR, NAME + "*.so"))
    if not cands:
        raise RuntimeError(f"torch extension not built in {BUILD_DIR} (run build_extension())")
    spec = importlib.util.spec_from_file_location(NAME, cands[0])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[NAME] = mod
    spec.loader.exec_module(mod)
    return mod


def dg_solve(log_kappa, go, reduction=1e-12, maxit=20000, precond=1):
    """u(log_kappa) for the grid operator ``go`` (a dune_pdelab_amd.GridOperator created with
    kappa_per_cell); differentiable with respect to log_kappa. Forward and adjoint solves are
    CG with ``precond`` (0 none, 1 point Jacobi, 2 block Jacobi)."""
    mod = load()
    return mod.dg_solve(log_kappa, int(go.h.value), reduction, maxit, precond)
