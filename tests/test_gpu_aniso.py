"""GPU parity of the cell-wise diagonal diffusion tensor A = diag(a_x, a_y, a_z)
(kappa_per_cell = 2, three values per cell).

* y += J z, y = J z, J^T z, the point diagonal and the FP32 apply against the oracle on
  operator variants (SIPG / NIPG / IIPG, weightsOff, convection with reaction, Q1-Q3) and on
  grids through both apply kernels (aligned rows for the row kernel, general grids): 1e-12
  relative (2e-5 for FP32);
* the residual of the interpolant of a quadratic with flux-continuous coefficients (the
  oracle's polynomial-reproduction pin) is zero on the GPU too;
* the assembled Jacobian and the block diagonal against the oracle's; CG with block Jacobi,
  the block-SOR sweep and the iterative block Jacobi against the oracle's;
* the rejections: a kappa array of the wrong length, the log-kappa set-up and the kappa
  gradient (isotropic only).
"""
import numpy as np
import pytest

from test_gpu_parity import APPLY_RTOL, rel, say

pytestmark = pytest.mark.gpu

ANISO = [
    ("sipg_q2_4x3x3", dict(nx=4, ny=3, nz=3, degree=2, alpha=3.0)),
    ("sipg_q3_3x2x3", dict(nx=3, ny=2, nz=3, degree=3, alpha=3.0)),
    ("nipg_q1_5x3x2", dict(nx=5, ny=3, nz=2, degree=1, method=0, alpha=2.0)),
    ("iipg_q2_off", dict(nx=3, ny=3, nz=2, degree=2, method=2, weights=1, alpha=4.0)),
    ("conv_react_q3", dict(nx=2, ny=3, nz=3, degree=3, alpha=3.0, b=(1.0, -0.5, 0.25), c=0.4, kappa0=0.1)),
    ("outflow_neumann_q2", dict(nx=3, ny=2, nz=3, degree=2, alpha=3.0, b=(1.0, 0.0, 0.0), bctype=(0, -2, -1, 0, 0, -1))),
    # aligned rows: the row kernel (16 cells Q3, 32 Q2, 64 Q1)
    ("row_q3_32x2x3", dict(nx=32, ny=2, nz=3, degree=3, alpha=3.0)),
    ("row_q2_32x3x2_conv", dict(nx=32, ny=3, nz=2, degree=2, alpha=3.0, b=(0.5, 1.0, -0.25))),
    ("row_q1_64x2x2", dict(nx=64, ny=2, nz=2, degree=1, alpha=3.0)),
]


def make_aniso(pkg, of, kw, seed, problem=None):
    kw = dict(kw, kappa_per_cell=2)
    if problem is not None:
        kw["problem"] = problem
    cfg = pkg.make_config(**kw)
    a = np.exp(np.random.default_rng(seed).normal(0.0, 1.0, (pkg.ncells(cfg), 3))).ravel()
    return cfg, a, pkg.GridOperator(cfg, a), of.Oracle(cfg, a)


@pytest.mark.parametrize("name,kw", ANISO, ids=[c[0] for c in ANISO])
def test_aniso_apply_vs_oracle(pkg, of, name, kw):
    say(f"[gpu] anisotropic apply {name}")
    cfg, a, go, o = make_aniso(pkg, of, kw, 71, problem=pkg.PROBLEM_EXP)
    rng = np.random.default_rng(72)
    z, y0 = rng.uniform(-1, 1, go.N), rng.uniform(-1, 1, go.N)
    zd, yd = go.vector(z), go.vector(y0)
    go.jacobian_apply(zd, yd)
    assert rel(yd.download(), o.jacobian_apply(z, y0.copy())) <= APPLY_RTOL
    go.operator_apply(zd, yd)
    assert rel(yd.download(), o.jacobian_apply(z)) <= APPLY_RTOL
    yd.zero()
    go.jacobian_apply_transpose(zd, yd)
    assert rel(yd.download(), o.jacobian_apply_transpose(z)) <= APPLY_RTOL
    dd = go.vector()
    go.point_diagonal(dd)
    assert rel(dd.download(), o.point_diagonal()) <= APPLY_RTOL
    rd = go.vector()
    go.residual(zd, rd)
    assert rel(rd.download(), o.residual(z)) <= APPLY_RTOL
    z32 = go.vector(z.astype(np.float32), dtype=np.float32)
    y32 = go.vector(dtype=np.float32)
    go.jacobian_apply(z32, y32)
    assert rel(y32.download().astype(np.float64), o.jacobian_apply(z)) <= 2e-5
    go.close()


def test_aniso_polynomial_reproduction(pkg, of):
    """The interpolant of a quadratic with a_x depending on the cell's (y, z) index, a_y on
    (x, z), a_z on (x, y) (continuous normal fluxes) has zero GPU residual."""
    say("[gpu] anisotropic polynomial reproduction")
    q = (0.3, 1.0, 2.0, -1.0, 1.0, 1.5, -2.0)
    nx, ny, nz = 4, 3, 3
    rng = np.random.default_rng(17)
    ax, ay, az = (np.exp(rng.normal(size=s)) for s in ((nz, ny), (nz, nx), (ny, nx)))
    a = np.array([(ax[k, j], ay[k, i], az[j, i]) for k in range(nz) for j in range(ny) for i in range(nx)])
    cfg = pkg.make_config(nx=nx, ny=ny, nz=nz, degree=2, alpha=4.0, kappa_per_cell=2,
                          problem=pkg.PROBLEM_QUADRATIC, quad=q, b=(0.5, -0.25, 0.75), c=0.3)
    go, o = pkg.GridOperator(cfg, a.ravel()), of.Oracle(cfg, a.ravel())
    src = -2.0 * ((a - 1.0) @ np.array(q[4:]))
    go.set_cell_source(src)
    o.set_cell_source(src)
    u = o.interpolate()
    rd = go.vector()
    go.residual(go.vector(u), rd)
    r = rd.download()
    assert np.abs(r).max() < 1e-11
    go.close()


@pytest.mark.parametrize("name,kw", [ANISO[0], ANISO[4], ANISO[6]], ids=[ANISO[i][0] for i in (0, 4, 6)])
def test_aniso_blocks_and_preconditioners(pkg, of, name, kw):
    say(f"[gpu] anisotropic blocks and preconditioners {name}")
    cfg, a, go, o = make_aniso(pkg, of, kw, 73, problem=pkg.PROBLEM_EXP)
    ref_blocks, ref_cols = o.assemble()
    bd, cd = go.jacobian()
    assert (cd.download().view(np.int64).reshape(ref_cols.shape) == ref_cols).all()
    assert rel(bd.download().reshape(ref_blocks.shape), ref_blocks) <= 1e-12
    nloc = (cfg.degree + 1) ** 3
    bdiag = pkg.DeviceVector(go, go.ncells * nloc * nloc)
    go.block_diagonal(bdiag)
    assert rel(bdiag.download().reshape(-1, nloc, nloc), o.block_diagonal()) <= 1e-12
    d = np.random.default_rng(74).uniform(-1, 1, go.N)
    v = go.vector()
    v.zero()
    go.block_sor_apply(go.vector(d), v)
    assert rel(v.download(), o.block_sor_apply(d)) <= 1e-10
    v.zero()
    its = go.iterative_block_jacobi_apply(go.vector(d), v)
    want, its_ref = o.iterative_block_jacobi_apply(d)
    assert rel(v.download(), want) <= 1e-4
    assert abs(its - its_ref) <= max(2, its_ref // 50)
    method = 0 if cfg.b[0] == 0.0 and cfg.b[1] == 0.0 and cfg.b[2] == 0.0 else 1
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=method, precond=pkg.PREC_BLOCK_JACOBI, reduction=1e-10)
    x_ref, res_ref = o.linear_problem_solve(method=method, precond=pkg.PREC_BLOCK_JACOBI, reduction=1e-10)
    assert res["solver"]["converged"] and res_ref["solver"]["converged"]
    it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
    assert abs(it - it_ref) <= (2 if method == 0 else max(2, 0.15 * it_ref))
    assert np.linalg.norm(o.residual(xd.download())) <= 1.01e-10 * res["first_defect"]
    go.close()


def test_aniso_rejections(pkg):
    say("[gpu] anisotropic rejections")
    cfg = pkg.make_config(n=2, degree=1, kappa_per_cell=2)
    with pytest.raises(ValueError):
        pkg.GridOperator(cfg, np.ones(8))
    go = pkg.GridOperator(cfg, np.ones(24))
    with pytest.raises(ValueError):
        go.set_kappa(np.ones(8))
    go.set_kappa(np.full(24, 2.0))
    with pytest.raises(RuntimeError, match="isotropic"):
        go.set_log_kappa_device(go.cell_vector())
    u = go.vector(np.ones(go.N))
    with pytest.raises(RuntimeError, match="isotropic"):
        go.grad_logkappa(u, u, go.cell_vector())
    bad = np.ones(24)
    bad[5] = -1.0
    with pytest.raises(RuntimeError):
        go.set_kappa(bad)
    go.close()
