"""ctypes mirror of ``pdelab_config`` and the result structs of include/pdelab_gpu.h.

The field order must match the C header exactly; ``tests/test_boundary.py`` checks
the struct size against the C compiler's.
"""
import ctypes

import numpy as np

ABI_VERSION = 2
MAX_GAUSSIANS = 16

# ConvectionDiffusionDGMethod / Weights / BoundaryConditions
# (R:dune/pdelab/localoperator/convectiondiffusiondg.hh:29-37,
#  R:dune/pdelab/localoperator/convectiondiffusionparameter.hh:111-115)
NIPG, SIPG, IIPG = 0, 1, 2
WEIGHTS_ON, WEIGHTS_OFF = 0, 1
BC_DIRICHLET, BC_NEUMANN, BC_OUTFLOW, BC_NONE = 1, -1, -2, -3

PROBLEM_ZERO, PROBLEM_EXP, PROBLEM_CONSTANT, PROBLEM_GAUSSIANS, PROBLEM_QUADRATIC, PROBLEM_TABULATED = range(6)
SOLVER_CG, SOLVER_BICGSTAB = 0, 1
PREC_NONE, PREC_JACOBI, PREC_BLOCK_JACOBI, PREC_BLOCK_SOR, PREC_ITERATIVE_BLOCK_JACOBI = 0, 1, 2, 3, 4
# variants of the mixed-precision solve, OR'd into its method (PDELAB_MIXED_RESTARTED / _FLEXIBLE)
MIXED_RESTARTED, MIXED_FLEXIBLE = 16, 32

STATUS = {
    0: "PDELAB_OK",
    1: "PDELAB_ERR_INVALID_ARG",
    2: "PDELAB_ERR_LINEARITY_MISMATCH",
    3: "PDELAB_ERR_HIP",
    4: "PDELAB_ERR_UNSUPPORTED",
    5: "PDELAB_ERR_SOLVER_ABORT",
    6: "PDELAB_ERR_OUTFLOW_ON_INFLOW",
    7: "PDELAB_ERR_NO_DEVICE",
}


class PdelabConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("degree", ctypes.c_int),
        ("method", ctypes.c_int),
        ("weights", ctypes.c_int),
        ("alpha", ctypes.c_double),
        ("intorderadd", ctypes.c_int),
        ("kappa_per_cell", ctypes.c_int),
        ("kappa0", ctypes.c_double),
        ("b", ctypes.c_double * 3),
        ("c", ctypes.c_double),
        ("bctype", ctypes.c_int * 6),
        ("problem", ctypes.c_int),
        ("f0", ctypes.c_double), ("g0", ctypes.c_double),
        ("quad", ctypes.c_double * 7),
        ("n_gaussians", ctypes.c_int),
        ("gaussians", (ctypes.c_double * 5) * MAX_GAUSSIANS),
        ("j0", ctypes.c_double), ("o0", ctypes.c_double),
    ]


class LinearSolverResult(ctypes.Structure):
    """LinearSolverResult<double>, R:dune/pdelab/backend/solver.hh:29-51"""
    _fields_ = [
        ("converged", ctypes.c_int),
        ("iterations", ctypes.c_uint),
        ("elapsed", ctypes.c_double),
        ("reduction", ctypes.c_double),
        ("conv_rate", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LinearProblemResult(ctypes.Structure):
    """StationaryLinearProblemSolverResult, R:dune/pdelab/stationary/linearproblem.hh:27-44"""
    _fields_ = [
        ("solver", LinearSolverResult),
        ("first_defect", ctypes.c_double),
        ("defect", ctypes.c_double),
        ("assembler_time", ctypes.c_double),
        ("linear_solver_time", ctypes.c_double),
        ("linear_solver_iterations", ctypes.c_int),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "solver"}
        d["solver"] = self.solver.as_dict()
        return d


TS_IMPLICIT_EULER, TS_CRANK_NICOLSON, TS_ALEXANDER2 = 0, 1, 2


class OneStepResult(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("stages", ctypes.c_int),
        ("linear_solver_iterations", ctypes.c_int),
        ("linear_solver_time", ctypes.c_double),
        ("elapsed", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MixedResult(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int),
        ("outer_iterations", ctypes.c_int),
        ("inner_iterations", ctypes.c_int),
        ("first_defect", ctypes.c_double),
        ("defect", ctypes.c_double),
        ("reduction", ctypes.c_double),
        ("elapsed", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def make_config(n=4, degree=1, nx=None, ny=None, nz=None, lx=1.0, ly=1.0, lz=1.0,
                method=SIPG, weights=WEIGHTS_ON, alpha=1.0, intorderadd=0,
                kappa_per_cell=False, kappa0=1.0, b=(0.0, 0.0, 0.0), c=0.0,
                bctype=(0, 0, 0, 0, 0, 0), problem=PROBLEM_ZERO, f0=0.0, g0=0.0,
                quad=(0.0,) * 7, gaussians=(), j0=0.0, o0=0.0):
    """Config with the reference defaults (SIPG, weightsOn, alpha=1,
    convectiondiffusiondg.hh:85-90), overridable by keyword."""
    cfg = PdelabConfig()
    cfg.abi_version = ABI_VERSION
    cfg.nx = nx if nx is not None else n
    cfg.ny = ny if ny is not None else n
    cfg.nz = nz if nz is not None else n
    cfg.lx, cfg.ly, cfg.lz = lx, ly, lz
    cfg.degree = degree
    cfg.method = method
    cfg.weights = weights
    cfg.alpha = alpha
    cfg.intorderadd = intorderadd
    # False/True, or 2: a diagonal tensor diag(a_x, a_y, a_z) per cell (three values per cell)
    cfg.kappa_per_cell = 2 if kappa_per_cell == 2 else (1 if kappa_per_cell else 0)
    cfg.kappa0 = kappa0
    for i in range(3):
        cfg.b[i] = b[i]
    cfg.c = c
    for i in range(6):
        cfg.bctype[i] = bctype[i]
    cfg.problem = problem
    cfg.f0, cfg.g0 = f0, g0
    for i in range(7):
        cfg.quad[i] = quad[i]
    if len(gaussians) > MAX_GAUSSIANS:
        raise ValueError("too many gaussians")
    cfg.j0, cfg.o0 = j0, o0
    cfg.n_gaussians = len(gaussians)
    for i, gsn in enumerate(gaussians):
        for j in range(5):
            cfg.gaussians[i][j] = gsn[j]
    return cfg


def ncells(cfg):
    return cfg.nx * cfg.ny * cfg.nz


def ndofs(cfg):
    return ncells(cfg) * (cfg.degree + 1) ** 3


def quadrature_points(cfg):
    """Gauss points per direction of the operator's quadrature (order 2k + intorderadd)."""
    return (2 * cfg.degree + cfg.intorderadd) // 2 + 1


def tabulate_velocity(cfg, b):
    """The arrays pdelab_gpu_set_velocity_field takes, from a velocity field b(x, y, z) that
    accepts numpy arrays and returns (b_x, b_y, b_z):
    b_vol  per cell (x fastest), the m^3 volume Gauss points (x fastest), (b_x, b_y, b_z);
    b_face per face normal to x, then y, then z (within a direction lexicographic in the
           face position, x fastest, the position along the normal running 0..n), the m^2
           face points (the lower tangential direction fastest), the normal component b_d.
    The points are the Gauss-Legendre points of quadratureRule(geo, 2k + intorderadd) mapped
    to each cell. b is evaluated at the physical point, so on an interior face both cells
    see the same value (a continuous field)."""
    m = quadrature_points(cfg)
    xi, _ = np.polynomial.legendre.leggauss(m)
    xi = (xi + 1.0) / 2.0
    n = (cfg.nx, cfg.ny, cfg.nz)
    h = (cfg.lx / cfg.nx, cfg.ly / cfg.ny, cfg.lz / cfg.nz)
    # volume: cell (i, j, k), point (a, b, c) -> coordinates
    ci = [np.arange(n[d]) for d in range(3)]
    X = (ci[0][None, None, :, None, None, None] + xi[None, None, None, None, None, :]) * h[0]
    Y = (ci[1][None, :, None, None, None, None] + xi[None, None, None, None, :, None]) * h[1]
    Z = (ci[2][:, None, None, None, None, None] + xi[None, None, None, :, None, None]) * h[2]
    shape = (n[2], n[1], n[0], m, m, m)
    bv = [np.broadcast_to(np.asarray(v, dtype=np.float64), shape)
          for v in np.broadcast_arrays(*b(*np.broadcast_arrays(X, Y, Z)))]
    b_vol = np.stack(bv, axis=-1).reshape(-1)
    faces = []
    for d in range(3):
        t1, t2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
        npos = [n[e] + (1 if e == d else 0) for e in range(3)]
        pos = [np.arange(npos[e]) for e in range(3)]
        coord = [None, None, None]
        # face (z, y, x) positions, then points (q2 along t2, q1 along t1)
        for e in range(3):
            p = pos[e].astype(np.float64).reshape([-1 if a == 2 - e else 1 for a in range(3)] + [1, 1])
            if e == d:
                coord[e] = p * h[e]
            elif e == t1:
                coord[e] = (p + xi.reshape(1, 1, 1, 1, m)) * h[e]
            else:
                coord[e] = (p + xi.reshape(1, 1, 1, m, 1)) * h[e]
        bd = b(*np.broadcast_arrays(*coord))[d]
        faces.append(np.broadcast_to(np.asarray(bd, dtype=np.float64),
                                     (npos[2], npos[1], npos[0], m, m)).reshape(-1))
    return b_vol, np.concatenate(faces)
