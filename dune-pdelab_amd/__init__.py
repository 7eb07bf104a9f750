"""dune-pdelab_amd -- Python handle on the gfx950 matrix-free SIPG DG hot path.

This module is a thin ctypes layer over the C-ABI library ``lib/libpdelab_gpu.so``
(declared in include/pdelab_gpu.h); all numerics run in the HIP kernels. There is no
CPU fallback: if the library is missing or no HIP device is present, the calls
raise.

Import it under the name ``dune_pdelab_amd`` (the directory name has a hyphen)::

    spec = importlib.util.spec_from_file_location(
        "dune_pdelab_amd", "<repo>/dune-pdelab_amd/__init__.py",
        submodule_search_locations=["<repo>/dune-pdelab_amd"])
"""
import ctypes
import os
import subprocess
import weakref

import numpy as np

from . import config as cfgmod
from .config import (  # noqa: F401  (re-exported)
    BC_DIRICHLET, BC_NEUMANN, BC_NONE, BC_OUTFLOW, IIPG, MIXED_FLEXIBLE, MIXED_RESTARTED, NIPG, PREC_BLOCK_JACOBI,
    PREC_BLOCK_SOR, PREC_ITERATIVE_BLOCK_JACOBI, PREC_JACOBI, PREC_NONE,
    PROBLEM_CONSTANT, PROBLEM_EXP, PROBLEM_GAUSSIANS, PROBLEM_QUADRATIC, PROBLEM_TABULATED,
    PROBLEM_ZERO, SIPG, SOLVER_BICGSTAB, SOLVER_CG, TS_ALEXANDER2, TS_CRANK_NICOLSON, TS_IMPLICIT_EULER,
    WEIGHTS_OFF, WEIGHTS_ON, LinearProblemResult, LinearSolverResult, MixedResult, OneStepResult, PdelabConfig,
    make_config, ncells, ndofs, quadrature_points, tabulate_velocity)

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libpdelab_gpu.so")

# cells of the largest grid pdelab_gpu_l2_error_squared_noalloc takes (PDELAB_NOALLOC_MAX_CELLS)
NOALLOC_MAX_CELLS = 1048576

_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


class PdelabError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{cfgmod.STATUS.get(status, status)}: {message}")
        self.status = status


def build():
    """Compiles lib/libpdelab_gpu.so for gfx950 (hipcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-j4", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"HIP library not built: {LIBPATH} (run __graft_entry__.build())")
    L = ctypes.CDLL(LIBPATH)
    C = ctypes.POINTER(cfgmod.PdelabConfig)
    L.pdelab_gpu_default_config.argtypes = [C]
    L.pdelab_gpu_default_config.restype = None
    L.pdelab_gpu_create.argtypes = [C, _dp, ctypes.POINTER(_vp)]
    L.pdelab_gpu_destroy.argtypes = [_vp]
    L.pdelab_gpu_last_error.argtypes = [_vp]
    L.pdelab_gpu_last_error.restype = ctypes.c_char_p
    for fn in ("pdelab_gpu_size", "pdelab_gpu_num_cells"):
        getattr(L, fn).argtypes = [_vp]
        getattr(L, fn).restype = ctypes.c_size_t
    L.pdelab_gpu_is_linear.argtypes = [_vp]
    L.pdelab_gpu_quadrature_points.argtypes = [_vp]
    L.pdelab_gpu_set_kappa.argtypes = [_vp, _dp]
    L.pdelab_gpu_set_log_kappa_device.argtypes = [_vp, _vp]
    L.pdelab_gpu_set_cell_source.argtypes = [_vp, _dp]
    L.pdelab_gpu_set_tabulated_data.argtypes = [_vp, _dp, _dp]
    L.pdelab_gpu_set_velocity_field.argtypes = [_vp, _dp, _dp]
    L.pdelab_gpu_malloc.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_size_t]
    L.pdelab_gpu_free.argtypes = [_vp, _vp]
    for fn in ("pdelab_gpu_memcpy_htod", "pdelab_gpu_memcpy_dtoh", "pdelab_gpu_memcpy_dtod"):
        getattr(L, fn).argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
    L.pdelab_gpu_memset_zero.argtypes = [_vp, _vp, ctypes.c_size_t]
    L.pdelab_gpu_synchronize.argtypes = [_vp]
    for fn in ("pdelab_gpu_residual", "pdelab_gpu_jacobian_apply", "pdelab_gpu_jacobian_apply_f32",
               "pdelab_gpu_jacobian_apply_transpose", "pdelab_gpu_mass_apply", "pdelab_gpu_grad_source"):
        getattr(L, fn).argtypes = [_vp, _vp, _vp]
    L.pdelab_gpu_jacobian_apply_nonlinear.argtypes = [_vp, _vp, _vp, _vp]
    L.pdelab_gpu_point_diagonal.argtypes = [_vp, _vp]
    L.pdelab_gpu_operator_apply.argtypes = [_vp, _vp, _vp]
    L.pdelab_gpu_block_diagonal.argtypes = [_vp, ctypes.c_int, _vp]
    L.pdelab_gpu_jacobian.argtypes = [_vp, ctypes.c_int, _vp, _vp]
    L.pdelab_gpu_l2_error_squared.argtypes = [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.pdelab_gpu_l2_error_squared_noalloc.argtypes = [_vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.pdelab_gpu_sample.argtypes = [_vp, _vp, ctypes.c_int, _vp]
    L.pdelab_gpu_onestep_jacobian_apply.argtypes = [_vp, ctypes.c_double, ctypes.c_double, _vp, _vp]
    L.pdelab_gpu_onestep_step.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, _vp, _vp,
                                          ctypes.POINTER(cfgmod.OneStepResult)]
    L.pdelab_gpu_block_matrix_apply.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pdelab_gpu_block_jacobi_apply.argtypes = [_vp, _vp, _vp]
    L.pdelab_gpu_block_jacobi_apply_f32.argtypes = [_vp, _vp, _vp]
    L.pdelab_gpu_block_sor_apply.argtypes = [_vp, _vp, _vp]
    L.pdelab_gpu_set_sor_relaxation.argtypes = [_vp, ctypes.c_double]
    L.pdelab_gpu_iterative_block_jacobi_apply.argtypes = [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_ulonglong)]
    L.pdelab_gpu_set_block_solver_options.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.pdelab_gpu_dot.argtypes = [_vp, _vp, _vp, _dp]
    L.pdelab_gpu_norm.argtypes = [_vp, _vp, _dp]
    L.pdelab_gpu_axpy.argtypes = [_vp, ctypes.c_double, _vp, _vp]
    for fn in ("pdelab_gpu_solve", "pdelab_gpu_solve_transpose"):
        getattr(L, fn).argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp,
                                   ctypes.POINTER(cfgmod.LinearSolverResult)]
    L.pdelab_gpu_linear_problem_solve.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, _vp,
                                                  ctypes.POINTER(cfgmod.LinearProblemResult)]
    L.pdelab_gpu_solve_mixed.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, _vp, ctypes.POINTER(cfgmod.MixedResult)]
    L.pdelab_gpu_grad_logkappa.argtypes = [_vp, _vp, _vp, _vp]
    L.pdelab_gpu_time_apply.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp]
    L.pdelab_gpu_time_stream.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _dp]
    _lib = L
    return L


class DeviceVector:
    """A device buffer owned by a GridOperator handle (float64 or float32)."""

    def __init__(self, go, n, dtype=np.float64):
        self.go = go
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        ptr = _vp()
        go._check(lib().pdelab_gpu_malloc(go.h, ctypes.byref(ptr), self.nbytes))
        self.ptr = ptr
        go._children.add(self)  # close() frees the buffers still alive
        self.zero()

    @property
    def p(self):
        return self.ptr

    def zero(self):
        self.go._check(lib().pdelab_gpu_memset_zero(self.go.h, self.ptr, self.nbytes))
        return self

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n, (a.size, self.n)
        self.go._check(lib().pdelab_gpu_memcpy_htod(self.go.h, self.ptr, a.ctypes.data_as(_vp), self.nbytes))
        return self

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.go._check(lib().pdelab_gpu_memcpy_dtoh(self.go.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        self.go._check(lib().pdelab_gpu_memcpy_dtod(self.go.h, self.ptr, other.ptr, self.nbytes))
        return self

    def free(self):
        if getattr(self, "ptr", None) is not None and self.go.h:
            lib().pdelab_gpu_free(self.go.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GridOperator:
    """GridOperator + ISTL-backend handle of the HIP implementation.

    Mirrors GridOperator::residual / jacobian_apply
    (R:dune/pdelab/gridoperator/gridoperator.hh:176-205) and the solver backends'
    apply(z, r, reduction) (R:dune/pdelab/backend/istl/seqistlsolverbackend.hh:231-252).
    """

    def __init__(self, cfg, kappa=None):
        L = lib()
        self.cfg = cfg
        self._children = weakref.WeakSet()  # live DeviceVectors of this handle
        kp = None
        if kappa is not None:
            self._kappa = np.ascontiguousarray(kappa, dtype=np.float64)
            if cfg.kappa_per_cell and self._kappa.size != self._kappa_size(cfg):
                raise ValueError(f"kappa needs {self._kappa_size(cfg)} values, got {self._kappa.size}")
            kp = self._kappa.ctypes.data_as(_dp)
        h = _vp()
        st = L.pdelab_gpu_create(ctypes.byref(cfg), kp, ctypes.byref(h))
        if st != 0:
            raise PdelabError(st, L.pdelab_gpu_last_error(None).decode())
        self.h = h
        self.N = L.pdelab_gpu_size(h)
        self.ncells = L.pdelab_gpu_num_cells(h)

    def close(self):
        """Frees the device buffers of the vectors still alive, then the handle
        (pdelab_gpu_destroy does not free user allocations)."""
        if getattr(self, "h", None):
            for v in list(self._children):
                v.free()
            lib().pdelab_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise PdelabError(st, lib().pdelab_gpu_last_error(self.h).decode())

    def _vec(self, v, n=None, dtype=np.float64):
        """Device pointer of v after checking that it is a live DeviceVector of n elements
        (default: the DoF count) and the given dtype. The library takes raw pointers and
        cannot check them; a shorter or narrower buffer would be read and written past its
        end."""
        if not isinstance(v, DeviceVector) or v.ptr is None:
            raise TypeError("expected a live DeviceVector")
        if v.dtype != np.dtype(dtype):
            raise TypeError(f"expected a {np.dtype(dtype)} vector, got {v.dtype}")
        want = self.N if n is None else n
        if v.n != want:
            raise ValueError(f"expected {want} elements, got {v.n}")
        return v.ptr

    def _pair(self, z, y):
        """Pointers of two DoF vectors of one dtype (float64 or float32) and that dtype."""
        dt = z.dtype if isinstance(z, DeviceVector) else np.dtype(np.float64)
        if dt not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise TypeError(f"unsupported vector dtype {dt}")
        return self._vec(z, dtype=dt), self._vec(y, dtype=dt), dt

    # -- vectors
    def vector(self, data=None, dtype=np.float64):
        v = DeviceVector(self, self.N, dtype)
        if data is not None:
            v.upload(data)
        return v

    def cell_vector(self, data=None):
        v = DeviceVector(self, self.ncells, np.float64)
        if data is not None:
            v.upload(data)
        return v

    def synchronize(self):
        self._check(lib().pdelab_gpu_synchronize(self.h))

    # -- parameters
    @staticmethod
    def _kappa_size(cfg):
        # one value per cell, or a_x, a_y, a_z per cell for kappa_per_cell == 2
        return cfgmod.ncells(cfg) * (3 if cfg.kappa_per_cell == 2 else 1)

    def set_kappa(self, kappa):
        k = np.ascontiguousarray(kappa, dtype=np.float64)
        if k.size != self._kappa_size(self.cfg):
            raise ValueError(f"kappa needs {self._kappa_size(self.cfg)} values, got {k.size}")
        self._check(lib().pdelab_gpu_set_kappa(self.h, k.ctypes.data_as(_dp)))

    def set_log_kappa_device(self, logk):
        self._check(lib().pdelab_gpu_set_log_kappa_device(self.h, self._vec(logk, self.ncells)))

    def set_cell_source(self, s):
        if s is None:
            self._check(lib().pdelab_gpu_set_cell_source(self.h, None))
        else:
            s = np.ascontiguousarray(s, dtype=np.float64)
            self._check(lib().pdelab_gpu_set_cell_source(self.h, s.ctypes.data_as(_dp)))

    def set_tabulated(self, fq, gq):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        gq = np.ascontiguousarray(gq, dtype=np.float64)
        self._check(lib().pdelab_gpu_set_tabulated_data(self.h, fq.ctypes.data_as(_dp), gq.ctypes.data_as(_dp)))

    def quadrature_points(self):
        return lib().pdelab_gpu_quadrature_points(self.h)

    def set_velocity_field(self, b_vol, b_face=None):
        """Velocity field b(x) at the quadrature points (layouts of tabulate_velocity); None
        restores the constant b of the configuration."""
        if b_vol is None:
            self._check(lib().pdelab_gpu_set_velocity_field(self.h, None, None))
            return
        m = self.quadrature_points()
        c = self.cfg
        nv = ncells(c) * m ** 3 * 3
        nf = ((c.nx + 1) * c.ny * c.nz + c.nx * (c.ny + 1) * c.nz + c.nx * c.ny * (c.nz + 1)) * m * m
        bv = np.ascontiguousarray(b_vol, dtype=np.float64).ravel()
        bf = np.ascontiguousarray(b_face, dtype=np.float64).ravel()
        if bv.size != nv or bf.size != nf:
            raise ValueError(f"velocity field: b_vol needs {nv} values and b_face {nf}")
        self._check(lib().pdelab_gpu_set_velocity_field(self.h, bv.ctypes.data_as(_dp), bf.ctypes.data_as(_dp)))

    # -- GridOperator
    def residual(self, x, r):
        self._check(lib().pdelab_gpu_residual(self.h, self._vec(x), self._vec(r)))

    def jacobian_apply(self, z, y):
        zp, yp, dt = self._pair(z, y)
        if dt == np.float32:
            self._check(lib().pdelab_gpu_jacobian_apply_f32(self.h, zp, yp))
        else:
            self._check(lib().pdelab_gpu_jacobian_apply(self.h, zp, yp))

    def operator_apply(self, z, y):
        """y = J z (OnTheFlyOperator::apply: y = 0, then jacobian_apply)."""
        self._check(lib().pdelab_gpu_operator_apply(self.h, self._vec(z), self._vec(y)))

    def jacobian_apply_nonlinear(self, x, z, y):
        self._check(lib().pdelab_gpu_jacobian_apply_nonlinear(self.h, self._vec(x), self._vec(z), self._vec(y)))

    def jacobian_apply_transpose(self, z, y):
        self._check(lib().pdelab_gpu_jacobian_apply_transpose(self.h, self._vec(z), self._vec(y)))

    def point_diagonal(self, d):
        self._check(lib().pdelab_gpu_point_diagonal(self.h, self._vec(d)))

    def block_diagonal(self, blocks, transpose=False):
        """blocks (ncells * nloc^2 doubles) = the diagonal blocks of J (J^T), row-major per cell."""
        self._check(lib().pdelab_gpu_block_diagonal(self.h, 1 if transpose else 0,
                                                    self._vec(blocks, self.ncells * ((self.cfg.degree + 1) ** 6))))

    def jacobian(self, transpose=False):
        """Assembled Jacobian (7-point block pattern) as device vectors (blocks, cols); cols holds
        int64 values in a float64-sized buffer (download with .download().view(np.int64))."""
        nloc = (self.cfg.degree + 1) ** 3
        blocks = DeviceVector(self, 7 * self.ncells * nloc * nloc)
        cols = DeviceVector(self, 7 * self.ncells)
        self._check(lib().pdelab_gpu_jacobian(self.h, 1 if transpose else 0, blocks.p, cols.p))
        return blocks, cols

    def block_matrix_apply(self, blocks, cols, z, y):
        """y = A z for a block matrix from jacobian()."""
        self._check(lib().pdelab_gpu_block_matrix_apply(self.h, self._vec(blocks, 7 * self.ncells * ((self.cfg.degree + 1) ** 6)),
                                                        self._vec(cols, 7 * self.ncells), self._vec(z),
                                                        self._vec(y)))

    # -- output
    def l2_error_squared(self, x, order=10):
        """integral of (u_h - u)^2 for the built-in problem's exact solution u."""
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_l2_error_squared(self.h, self._vec(x), order, ctypes.byref(out)))
        return out.value

    def l2_error_squared_noalloc(self, x, order=10):
        """l2_error_squared without any device allocation (library-owned scratch of
        NOALLOC_MAX_CELLS + 1 doubles per GPU; larger grids raise)."""
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_l2_error_squared_noalloc(self.h, self._vec(x), order, ctypes.byref(out)))
        return out.value

    def sample(self, x, intervals):
        """u_h at the (intervals+1)^3 equidistant local points of every cell, shape (ncells, L+1, L+1, L+1)
        indexed [cell, z, y, x]."""
        L1 = intervals + 1
        v = DeviceVector(self, self.ncells * L1 ** 3)
        self._check(lib().pdelab_gpu_sample(self.h, self._vec(x), intervals, v.p))
        out = v.download().reshape(self.ncells, L1, L1, L1)
        v.free()
        return out

    # -- instationary
    def onestep_jacobian_apply(self, a_mass, b_dt, z, y):
        """y += b_dt J z + a_mass M z (a stage of OneStepGridOperator)."""
        self._check(lib().pdelab_gpu_onestep_jacobian_apply(self.h, a_mass, b_dt, self._vec(z), self._vec(y)))

    def onestep_step(self, xold, xnew, scheme=TS_ALEXANDER2, dt=0.1, method=SOLVER_CG, precond=PREC_JACOBI,
                     reduction=1e-10, maxit=5000):
        """One step xold -> xnew of OneStepMethod::apply (xnew holds the first stage's initial guess)."""
        res = cfgmod.OneStepResult()
        self._check(lib().pdelab_gpu_onestep_step(self.h, scheme, method, precond, dt, reduction, maxit,
                                                  self._vec(xold), self._vec(xnew), ctypes.byref(res)))
        return res.as_dict()

    def block_jacobi_apply(self, z, y):
        """y += D^{-1} z with D the block diagonal of J (FP32 for float32 vectors)."""
        zp, yp, dt = self._pair(z, y)
        if dt == np.float32:
            self._check(lib().pdelab_gpu_block_jacobi_apply_f32(self.h, zp, yp))
        else:
            self._check(lib().pdelab_gpu_block_jacobi_apply(self.h, zp, yp))

    def block_sor_apply(self, d, v):
        """One block-SOR step on v in place (FP64); on v = 0 the preconditioner v = W^{-1} d."""
        self._check(lib().pdelab_gpu_block_sor_apply(self.h, self._vec(d), self._vec(v)))

    def iterative_block_jacobi_apply(self, d, v):
        """v += W^{-1} d with per-cell BiCGSTAB block solves (FP64). Returns the block solvers'
        iteration count summed over the cells."""
        its = ctypes.c_ulonglong(0)
        self._check(lib().pdelab_gpu_iterative_block_jacobi_apply(self.h, self._vec(d), self._vec(v), ctypes.byref(its)))
        return its.value

    def set_block_solver_options(self, reduction=1e-5, maxiter=100, precondition=True, weight=1.0):
        """BlockSolverOptions of the iterative block Jacobi preconditioner."""
        self._check(lib().pdelab_gpu_set_block_solver_options(self.h, float(reduction), int(maxiter),
                                                               1 if precondition else 0, float(weight)))

    def set_sor_relaxation(self, omega):
        """Relaxation factor omega of the block-SOR preconditioner (default 1, in (0, 2))."""
        self._check(lib().pdelab_gpu_set_sor_relaxation(self.h, float(omega)))

    def mass_apply(self, z, y):
        self._check(lib().pdelab_gpu_mass_apply(self.h, self._vec(z), self._vec(y)))

    # -- vector backend
    def dot(self, x, y):
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_dot(self.h, self._vec(x), self._vec(y), ctypes.byref(out)))
        return out.value

    def norm(self, x):
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_norm(self.h, self._vec(x), ctypes.byref(out)))
        return out.value

    def axpy(self, a, x, y):
        self._check(lib().pdelab_gpu_axpy(self.h, a, self._vec(x), self._vec(y)))

    # -- solvers
    def solve(self, z, r, method=SOLVER_CG, precond=PREC_JACOBI, reduction=1e-10, maxit=5000, transpose=False):
        res = LinearSolverResult()
        fn = lib().pdelab_gpu_solve_transpose if transpose else lib().pdelab_gpu_solve
        self._check(fn(self.h, method, precond, reduction, maxit, self._vec(z), self._vec(r), ctypes.byref(res)))
        return res.as_dict()

    def linear_problem_solve(self, x, method=SOLVER_CG, precond=PREC_JACOBI, reduction=1e-10, min_defect=1e-99,
                             maxit=5000):
        res = LinearProblemResult()
        self._check(lib().pdelab_gpu_linear_problem_solve(self.h, method, precond, reduction, min_defect, maxit,
                                                          self._vec(x), ctypes.byref(res)))
        return res.as_dict()

    def solve_mixed(self, x, method=SOLVER_CG, precond=PREC_JACOBI, inner_reduction=None, inner_maxit=5000,
                    outer_reduction=1e-10, outer_maxit=50, variant=0):
        """FP32 Krylov inside FP64 refinement. variant: 0, MIXED_RESTARTED or MIXED_FLEXIBLE
        (OR'd into the method, see pdelab_gpu_solve_mixed). inner_reduction defaults to the
        FP32 residual reduction between FP64 residual replacements with reliable updates, 0.1
        for CG and 1e-2 for BiCGSTAB, to 1e-2 for the flexible outer iteration and to 1e-4 for
        restarted refinement (the reduction of each inner solve)."""
        if inner_reduction is None:
            if variant & MIXED_FLEXIBLE:
                inner_reduction = 1e-2
            elif variant & MIXED_RESTARTED:
                inner_reduction = 1e-4
            else:
                inner_reduction = 0.1 if method == SOLVER_CG else 1e-2
        res = MixedResult()
        self._check(lib().pdelab_gpu_solve_mixed(self.h, method | variant, precond, inner_reduction, inner_maxit,
                                                 outer_reduction, outer_maxit, self._vec(x), ctypes.byref(res)))
        return res.as_dict()

    # -- adjoint gradients
    def grad_logkappa(self, u, lam, g):
        self._check(lib().pdelab_gpu_grad_logkappa(self.h, self._vec(u), self._vec(lam), self._vec(g, self.ncells)))

    def grad_source(self, lam, g):
        self._check(lib().pdelab_gpu_grad_source(self.h, self._vec(lam), self._vec(g, self.ncells)))

    # -- measurement
    def time_stream(self, x, y, warmup=3, reps=20):
        """Average duration (ms) of y += 0.5 x over FP64 vectors of the handle's length."""
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_time_stream(self.h, self._vec(x), self._vec(y), warmup, reps, ctypes.byref(out)))
        return out.value

    def time_apply(self, z, y, warmup=3, reps=20):
        zp, yp, dt = self._pair(z, y)
        prec = 32 if dt == np.float32 else 64
        out = ctypes.c_double()
        self._check(lib().pdelab_gpu_time_apply(self.h, prec, zp, yp, warmup, reps, ctypes.byref(out)))
        return out.value
