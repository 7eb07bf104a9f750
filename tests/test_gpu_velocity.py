"""GPU parity of the velocity field b(x) tabulated at the quadrature points
(pdelab_gpu_set_velocity_field, through the C-ABI), against the oracle on the same field:

* y += J z, y = J z, y += J^T z, R(x) (load vector with the field: the inflow term of Dirichlet
  sides and the manufactured f and j of the exact-solution problems), the point diagonal: 1e-12
  relative;
  the FP32 apply: 2e-5. Fields: random values at every point (the operator is defined for
  any data), and a divergence-free linear field where an outflow side needs b.n >= 0. Q1-Q3,
  SIPG / NIPG, Dirichlet / Neumann / outflow / none sides, cell-wise kappa, reaction, the
  rule of order 2k + 2 (intorderadd 2), grids whose rows take the row kernel;
* zero GPU residual of the interpolant of a quadratic with the divergence-free field (the
  oracle's consistency pin);
* BiCGSTAB with point Jacobi against the oracle's solve (iterations within 15 %, solution at
  the solver tolerance), and the mixed-precision BiCGSTAB at the oracle's residual; the
  transposed solve J^T z = r checked with the oracle's J^T;
* the cell blocks with the field: the block diagonal of J and J^T against the oracle's
  (1e-12), the assembled Jacobian of J and J^T against the oracle's assembly (1e-12) and its
  block-matrix apply against the matrix-free apply, one block-SOR step (from zero and in place)
  against the oracle's, the mixed-precision BiCGSTAB with the FP32 sweep, BiCGSTAB with block Jacobi and with the iterative block Jacobi against the oracle's
  solves, one iterative block Jacobi application against the oracle's;
* the one-step stage operator a M + b dt J and two Alexander2 steps (BiCGSTAB, point and block
  Jacobi) against the oracle's;
* a 24^3 Q2 apply against the oracle's two-sided apply;
* the rejections: the configuration's constant b not zero, wrong array sizes, inflow on an
  outflow side, CG;
  clearing the field restores the b = 0 operator.
"""
import numpy as np
import pytest

from test_gpu_parity import APPLY_RTOL, rel, say

pytestmark = pytest.mark.gpu


def div_free(x, y, z):
    """Linear, divergence-free, b_x >= 0 on the unit cube."""
    return (0.5 - (y - 0.5) + 0.3 * z, 0.25 + (x - 0.5) - 0.2 * z, 0.2 + 0.5 * (x - 0.5) - 0.3 * (y - 0.5))


def random_field(pkg, cfg, seed):
    m = pkg.quadrature_points(cfg)
    rng = np.random.default_rng(seed)
    nf = (cfg.nx + 1) * cfg.ny * cfg.nz + cfg.nx * (cfg.ny + 1) * cfg.nz + cfg.nx * cfg.ny * (cfg.nz + 1)
    return rng.normal(size=pkg.ncells(cfg) * m ** 3 * 3), rng.normal(size=nf * m ** 2)


CASES = [
    # name, config, kappa seed, field ("random" or "divfree")
    ("q1_dirichlet", dict(nx=4, ny=3, nz=2, degree=1, alpha=3.0), None, "random"),
    ("q2_dirichlet", dict(nx=3, ny=4, nz=3, degree=2, alpha=3.0), None, "random"),
    ("q3_dirichlet", dict(nx=3, ny=2, nz=3, degree=3, alpha=3.0), None, "random"),
    ("q2_nipg_var_react", dict(nx=3, ny=3, nz=2, degree=2, method=0, alpha=2.0, c=0.4), 5, "random"),
    ("q3_var_neumann_none", dict(nx=2, ny=3, nz=3, degree=3, alpha=3.0, bctype=(0, -1, 0, -3, -1, 0)), 6, "random"),
    ("q2_outflow_neumann", dict(nx=4, ny=2, nz=3, degree=2, alpha=3.0, kappa0=0.1, bctype=(0, -2, -1, 0, 0, -1)),
     None, "divfree"),
    ("q1_intorderadd2", dict(nx=3, ny=3, nz=3, degree=1, alpha=3.0, intorderadd=2), None, "random"),
    ("q3_intorderadd2_var", dict(nx=2, ny=2, nz=3, degree=3, alpha=3.0, intorderadd=2), 8, "random"),
    ("q2_aniso_cells", dict(nx=3, ny=2, nz=2, degree=2, lx=1.0, ly=0.5, lz=2.0, alpha=3.0), None, "random"),
    # rows of whole chunks: the row kernel runs the diffusion part
    ("row_q3_16x2x4", dict(nx=16, ny=2, nz=4, degree=3, alpha=3.0), None, "random"),
    ("row_q2_32x3x2_var", dict(nx=32, ny=3, nz=2, degree=2, alpha=3.0), 9, "random"),
    ("single_cell_q3", dict(nx=1, ny=1, nz=1, degree=3, alpha=3.0), None, "random"),
]


def make_field(pkg, of, kw, seed, field, problem=None):
    kw = dict(kw)
    if problem is not None:
        kw["problem"] = problem
    cfg = pkg.make_config(**kw)
    kappa = None
    if seed is not None:
        kappa = np.exp(np.random.default_rng(seed).normal(0.0, 1.0, pkg.ncells(cfg)))
        cfg.kappa_per_cell = 1
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    bv, bf = random_field(pkg, cfg, 77) if field == "random" else pkg.tabulate_velocity(cfg, div_free)
    go.set_velocity_field(bv, bf)
    o.set_velocity_field(bv, bf)
    return cfg, go, o


@pytest.mark.parametrize("name,kw,seed,field", CASES, ids=[c[0] for c in CASES])
def test_velocity_apply_vs_oracle(pkg, of, name, kw, seed, field):
    say(f"[gpu] velocity field apply {name}")
    cfg, go, o = make_field(pkg, of, kw, seed, field, problem=pkg.PROBLEM_EXP)
    rng = np.random.default_rng(31)
    z, y0 = rng.uniform(-1, 1, go.N), rng.uniform(-1, 1, go.N)
    zd, yd = go.vector(z), go.vector(y0)
    go.jacobian_apply(zd, yd)
    assert rel(yd.download(), o.jacobian_apply(z, y0.copy())) <= APPLY_RTOL
    go.operator_apply(zd, yd)
    want = o.jacobian_apply(z)
    assert rel(yd.download(), want) <= APPLY_RTOL
    yd.upload(y0)
    go.jacobian_apply_transpose(zd, yd)
    assert rel(yd.download(), o.jacobian_apply_transpose(z) + y0) <= APPLY_RTOL
    rd = go.vector()
    go.residual(zd, rd)
    assert rel(rd.download(), o.residual(z)) <= APPLY_RTOL
    dd = go.vector()
    go.point_diagonal(dd)
    assert rel(dd.download(), o.point_diagonal()) <= APPLY_RTOL
    z32 = go.vector(z.astype(np.float32), dtype=np.float32)
    y32 = go.vector(dtype=np.float32)
    go.jacobian_apply(z32, y32)
    assert rel(y32.download().astype(np.float64), want) <= 2e-5
    go.close()


@pytest.mark.parametrize("k", [1, 2, 3])
def test_velocity_polynomial_reproduction(pkg, of, k):
    say(f"[gpu] velocity field polynomial reproduction Q{k}")
    q = (0.3, 1.0, 2.0, -1.0, 0.0, 0.0, 0.0) if k == 1 else (0.3, 1.0, 2.0, -1.0, 1.0, 1.5, -2.0)
    kw = dict(nx=3, ny=2, nz=3, degree=k, alpha=4.0, kappa0=0.7, c=0.3, quad=q, bctype=(0, -2, -1, 0, 0, -1))
    cfg, go, o = make_field(pkg, of, kw, None, "divfree", problem=pkg.PROBLEM_QUADRATIC)
    u = o.interpolate()
    rd = go.vector()
    go.residual(go.vector(u), rd)
    assert np.abs(rd.download()).max() < 1e-11
    go.close()


@pytest.mark.parametrize("name,kw,seed", [
    ("q2_8cube_divfree", dict(n=8, degree=2, alpha=3.0, kappa0=0.05, problem=2, f0=1.0), None),
    ("q3_5cube_var", dict(n=5, degree=3, alpha=3.0, kappa0=0.5, problem=2, f0=1.0), 3),
], ids=["q2_8cube_divfree", "q3_5cube_var"])
def test_velocity_bicgstab_vs_oracle(pkg, of, name, kw, seed):
    say(f"[gpu] velocity field BiCGSTAB {name}")
    cfg, go, o = make_field(pkg, of, kw, seed, "divfree")
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=1, precond=pkg.PREC_JACOBI, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=1, precond=pkg.PREC_JACOBI, reduction=1e-10)
    assert res_ref["solver"]["converged"]
    it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
    say(f"  iterations {it} (oracle {it_ref})")
    assert abs(it - it_ref) <= max(2, 0.15 * it_ref)
    x = xd.download()
    r0 = np.linalg.norm(o.residual(np.zeros(go.N)))
    assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * r0
    assert rel(x, x_ref) <= 1e-7
    xm = go.vector()
    rm = go.solve_mixed(xm, method=1, precond=pkg.PREC_JACOBI, inner_reduction=1e-4, outer_reduction=1e-10)
    assert rm["converged"], rm
    assert np.linalg.norm(o.residual(xm.download())) <= 1.01e-10 * r0
    # the adjoint system J^T z = r
    rhs = np.random.default_rng(10).normal(size=go.N)
    zt, rt = go.vector(), go.vector(rhs)
    rs = go.solve(zt, rt, method=1, reduction=1e-12, transpose=True)
    assert rs["converged"]
    assert rel(o.jacobian_apply_transpose(zt.download()), rhs) <= 2e-12
    go.close()


@pytest.mark.parametrize("name,kw,seed,field", [CASES[1], CASES[4], CASES[5]], ids=[CASES[i][0] for i in (1, 4, 5)])
def test_velocity_cell_blocks(pkg, of, name, kw, seed, field):
    say(f"[gpu] velocity field cell blocks {name}")
    cfg, go, o = make_field(pkg, of, dict(kw, kappa0=0.2), seed, field, problem=pkg.PROBLEM_EXP)
    nloc = (cfg.degree + 1) ** 3
    ref = o.block_diagonal()
    bd = pkg.DeviceVector(go, go.ncells * nloc * nloc)
    go.block_diagonal(bd)
    assert rel(bd.download().reshape(-1, nloc, nloc), ref) <= 1e-12
    go.block_diagonal(bd, transpose=True)
    assert rel(bd.download().reshape(-1, nloc, nloc), ref.transpose(0, 2, 1)) <= 1e-12
    ref_blocks, ref_cols = o.assemble()
    jb, jc = go.jacobian()
    assert (jc.download().view(np.int64).reshape(ref_cols.shape) == ref_cols).all()
    assert rel(jb.download().reshape(ref_blocks.shape), ref_blocks) <= 1e-12
    z = np.random.default_rng(31).uniform(-1, 1, go.N)
    ym, yf = go.vector(), go.vector()
    go.block_matrix_apply(jb, jc, go.vector(z), ym)
    go.operator_apply(go.vector(z), yf)
    assert rel(ym.download(), yf.download()) <= 1e-12
    jt, _ = go.jacobian(transpose=True)
    want = np.zeros_like(ref_blocks)
    want[:, 0] = np.transpose(ref_blocks[:, 0], (0, 2, 1))
    for c in range(go.ncells):
        for f in range(6):
            cn = ref_cols[c, 1 + f]
            if cn >= 0:
                want[c, 1 + f] = ref_blocks[cn, 1 + (f ^ 1)].T
    assert rel(jt.download().reshape(ref_blocks.shape), want) <= 1e-12
    d = np.random.default_rng(74).uniform(-1, 1, go.N)
    v0 = np.random.default_rng(75).normal(size=go.N)
    dd, v = go.vector(d), go.vector()
    for omega in (1.0, 1.3):
        go.set_sor_relaxation(omega)
        o.set_sor_omega(omega)
        v.zero()
        go.block_sor_apply(dd, v)
        assert rel(v.download(), o.block_sor_apply(d)) <= 1e-10
        v.upload(v0)
        go.block_sor_apply(dd, v)
        assert rel(v.download(), o.block_sor_apply(d, v0)) <= 1e-10
    go.set_sor_relaxation(1.0)
    o.set_sor_omega(1.0)
    v.zero()
    its = go.iterative_block_jacobi_apply(go.vector(d), v)
    want, its_ref = o.iterative_block_jacobi_apply(d)
    assert rel(v.download(), want) <= 1e-4
    assert abs(its - its_ref) <= max(2, its_ref // 50)
    o.set_threads(8)
    # block solves to 1e-10 in the solves: with the default 1e-5 the outer BiCGSTAB's count
    # follows where rounding stops each inexact block solve
    go.set_block_solver_options(reduction=1e-10, maxiter=200)
    o.set_block_solver(1e-10, 200, 1, 1.0)
    for prec in (pkg.PREC_BLOCK_JACOBI, pkg.PREC_BLOCK_SOR, pkg.PREC_ITERATIVE_BLOCK_JACOBI):
        xd = go.vector()
        res = go.linear_problem_solve(xd, method=1, precond=prec, reduction=1e-10)
        x_ref, res_ref = o.linear_problem_solve(method=1, precond=prec, reduction=1e-10)
        assert res["solver"]["converged"] and res_ref["solver"]["converged"]
        it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
        say(f"  precond {prec}: iterations {it} (oracle {it_ref})")
        assert abs(it - it_ref) <= max(2, 0.15 * it_ref)
        assert np.linalg.norm(o.residual(xd.download())) <= 1.01e-10 * res["first_defect"]
    # FP32 block SOR inside the mixed-precision BiCGSTAB
    xm = go.vector()
    rm = go.solve_mixed(xm, method=1, precond=pkg.PREC_BLOCK_SOR, inner_reduction=1e-4, outer_reduction=1e-10)
    assert rm["converged"], rm
    assert np.linalg.norm(o.residual(xm.download())) <= 1.01e-10 * np.linalg.norm(o.residual(np.zeros(go.N)))
    go.close()


def test_velocity_onestep(pkg, of):
    """The instationary stage operator a M + b dt J with the field, and two Alexander2 steps
    (BiCGSTAB with point and with block Jacobi), against the oracle's."""
    say("[gpu] velocity field one-step operator and Alexander2 steps")
    kw = dict(n=5, degree=2, alpha=3.0, kappa0=0.05, problem=2, f0=1.0)
    cfg, go, o = make_field(pkg, of, kw, None, "divfree")
    o.set_threads(8)
    rng = np.random.default_rng(43)
    z, y0 = rng.uniform(-1, 1, go.N), rng.uniform(-1, 1, go.N)
    for a_mass, b_dt in ((1.0, 0.05), (0.0, 2.0), (1.0, 1e-4)):
        yd = go.vector(y0)
        go.onestep_jacobian_apply(a_mass, b_dt, go.vector(z), yd)
        assert rel(yd.download(), o.onestep_jacobian_apply(a_mass, b_dt, z, y0.copy())) <= 1e-12, (a_mass, b_dt)
    x0 = rng.normal(size=go.N) * 0.1
    for precond in (pkg.PREC_JACOBI, pkg.PREC_BLOCK_JACOBI):
        xo, xn = go.vector(x0), go.vector(x0)
        x_ref = x0.copy()
        for _ in range(2):
            res = go.onestep_step(xo, xn, scheme=pkg.TS_ALEXANDER2, dt=0.05, method=pkg.SOLVER_BICGSTAB,
                                  precond=precond, reduction=1e-12)
            assert res["converged"] and res["stages"] == 2
            xo.copy_from(xn)
            x_ref, rr = o.onestep_step(x_ref, pkg.TS_ALEXANDER2, 0.05, xnew=x_ref, method=pkg.SOLVER_BICGSTAB,
                                       precond=precond, reduction=1e-12)
            assert rr["converged"]
        assert rel(xn.download(), x_ref) <= 1e-8
    go.close()


def test_velocity_apply_24cube_q2(pkg, of):
    say("[gpu] velocity field apply 24^3 Q2 against the two-sided oracle")
    kw = dict(n=24, degree=2, alpha=3.0, kappa0=0.01, bctype=(0, -2, 0, 0, 0, 0))
    cfg, go, o = make_field(pkg, of, kw, None, "divfree", problem=pkg.PROBLEM_EXP)
    z = np.random.default_rng(5).uniform(-1, 1, go.N)
    yd = go.vector()
    go.operator_apply(go.vector(z), yd)
    assert rel(yd.download(), o.jacobian_apply_twosided(z, 8)) <= APPLY_RTOL
    go.close()


def test_velocity_rejections(pkg, of):
    say("[gpu] velocity field rejections")
    cfg = pkg.make_config(n=3, degree=2, alpha=3.0, b=(1.0, 0.0, 0.0))
    go = pkg.GridOperator(cfg)
    bv, bf = pkg.tabulate_velocity(cfg, div_free)
    with pytest.raises(RuntimeError, match="constant b must be zero"):
        go.set_velocity_field(bv, bf)
    go.close()
    cfg = pkg.make_config(n=3, degree=2, alpha=3.0, bctype=(-2, 0, 0, 0, 0, 0))
    go = pkg.GridOperator(cfg)
    with pytest.raises(RuntimeError, match="Outflow"):
        go.set_velocity_field(*pkg.tabulate_velocity(cfg, div_free))  # b_x > 0: inflow on -x
    go.close()
    cfg = pkg.make_config(n=3, degree=2, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    bv, bf = pkg.tabulate_velocity(cfg, div_free)
    with pytest.raises(ValueError):
        go.set_velocity_field(bv[:-1], bf)
    go.set_velocity_field(bv, bf)
    with pytest.raises(RuntimeError, match="CG"):
        go.linear_problem_solve(go.vector(), method=0)
    # cleared: the b = 0 operator again
    go.set_velocity_field(None)
    x = np.random.default_rng(2).normal(size=go.N)
    yd = go.vector()
    go.operator_apply(go.vector(x), yd)
    assert rel(yd.download(), o.jacobian_apply(x)) <= APPLY_RTOL
    go.close()
