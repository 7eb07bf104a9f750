"""GPU parity: the HIP path (through the C-ABI) against the oracle on identical inputs.

Tolerances (FP64): operator apply within 1e-12 relative (2-norm), as the target
states; solves at the solver tolerance. FP32 apply: within 1e-5 relative.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

APPLY_RTOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [
    ("sipg_q1_4x3x2", dict(nx=4, ny=3, nz=2, degree=1, alpha=3.0), None),
    ("sipg_q2_5x4x3", dict(nx=5, ny=4, nz=3, degree=2, alpha=3.0), None),
    ("sipg_q3_3x3x4", dict(nx=3, ny=3, nz=4, degree=3, alpha=3.0), None),
    ("sipg_q2_var", dict(nx=4, ny=4, nz=4, degree=2, alpha=3.0), 1),
    ("sipg_q3_var", dict(nx=3, ny=4, nz=5, degree=3, alpha=3.0), 2),
    ("nipg_q2_var", dict(nx=3, ny=2, nz=4, degree=2, method=0, alpha=1.0), 3),
    ("iipg_q1_var", dict(nx=5, ny=2, nz=3, degree=1, method=2, alpha=5.0), 4),
    ("weights_off_q3", dict(nx=2, ny=3, nz=3, degree=3, weights=1, alpha=3.0), 5),
    ("aniso_q2", dict(nx=4, ny=3, nz=2, degree=2, lx=1.0, ly=0.5, lz=2.0, alpha=3.0), 6),
    ("conv_react_q2", dict(nx=4, ny=4, nz=3, degree=2, alpha=3.0, b=(1.0, 0.5, -0.25), c=0.7, kappa0=0.01), None),
    ("conv_q3_var", dict(nx=3, ny=3, nz=3, degree=3, alpha=3.0, b=(-0.3, 0.0, 2.0)), 7),
    ("outflow_q2", dict(nx=3, ny=2, nz=3, degree=2, alpha=3.0, b=(1.0, 0.0, 0.0), bctype=(0, -2, 0, -1, -3, 0)), None),
    ("single_cell_q3", dict(nx=1, ny=1, nz=1, degree=3, alpha=3.0), None),
    # nx a multiple of the workgroup chunk (16 cells Q3, 32 Q2, 64 Q1): full row chunks only
    ("row_q3_32x3x3_var_conv", dict(nx=32, ny=3, nz=3, degree=3, alpha=3.0, b=(0.5, -1.0, 0.25)), 12),
    ("row_q3_16x2x4", dict(nx=16, ny=2, nz=4, degree=3, alpha=3.0, c=0.3), None),
    ("row_q2_32x3x2_var", dict(nx=32, ny=3, nz=2, degree=2, alpha=3.0), 13),
    ("row_q2_64x2x2_nipg_conv", dict(nx=64, ny=2, nz=2, degree=2, method=0, alpha=2.0, b=(-1.0, 0.5, 0.0)), None),
    ("row_q1_64x3x2_var", dict(nx=64, ny=3, nz=2, degree=1, alpha=3.0, b=(0.0, 0.0, 1.0)), 14),
    ("row_q1_128x2x1_off", dict(nx=128, ny=2, nz=1, degree=1, weights=1, alpha=3.0), None),
    ("line_x_q2", dict(nx=7, ny=1, nz=1, degree=2, alpha=3.0), None),
    # aligned rows with nz a multiple of the z-walk (2 chunks per workgroup at Q3, 8 or 4 at Q2)
    ("walk_q2_32x2x4", dict(nx=32, ny=2, nz=4, degree=2, alpha=3.0), None),
    ("walk_q3_16x3x6_react", dict(nx=16, ny=3, nz=6, degree=3, alpha=3.0, c=0.5), None),
    # the bench's kernel instance (Q3, aligned rows, z-walk of 2, no reaction, y = J z)
    ("walk_q3_32x2x4", dict(nx=32, ny=2, nz=4, degree=3, alpha=3.0), None),
    ("walk8_q2_32x2x8_react", dict(nx=32, ny=2, nz=8, degree=2, alpha=3.0, c=0.4), None),
]


def make(pkg, of, kw, seed, problem=None, **extra):
    kw = dict(kw)
    kw.update(extra)
    if problem is not None:
        kw["problem"] = problem
    cfg = pkg.make_config(**kw)
    kappa = None
    if seed is not None:
        rng = np.random.default_rng(seed)
        kappa = np.exp(rng.normal(0.0, 1.0, pkg.ncells(cfg)))
        cfg.kappa_per_cell = 1
    return cfg, kappa, pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)


@pytest.mark.parametrize("name,kw,seed", CASES, ids=[c[0] for c in CASES])
def test_apply_diag_transpose_mass(pkg, of, name, kw, seed):
    say(f"[gpu] apply parity case={name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    rng = np.random.default_rng(42)
    z = rng.uniform(-1, 1, go.N)
    y0 = rng.uniform(-1, 1, go.N)
    zd, yd = go.vector(z), go.vector(y0)
    go.jacobian_apply(zd, yd)
    y_ref = o.jacobian_apply(z, y0.copy())
    assert rel(yd.download(), y_ref) <= APPLY_RTOL
    # OnTheFlyOperator::apply: y = J z, whatever y held
    yd.upload(y0)
    go.operator_apply(zd, yd)
    assert rel(yd.download(), o.jacobian_apply(z)) <= APPLY_RTOL
    # transpose
    yd.upload(y0)
    go.jacobian_apply_transpose(zd, yd)
    assert rel(yd.download(), y0 + o.jacobian_apply_transpose(z)) <= APPLY_RTOL
    # point diagonal (accumulates)
    dd = go.vector(np.ones(go.N))
    go.point_diagonal(dd)
    assert rel(dd.download(), 1.0 + o.point_diagonal()) <= APPLY_RTOL
    # mass
    yd.zero()
    go.mass_apply(zd, yd)
    assert rel(yd.download(), o.mass_apply(z)) <= APPLY_RTOL
    # FP32 apply
    z32 = go.vector(z.astype(np.float32), dtype=np.float32)
    y32 = go.vector(np.zeros(go.N, np.float32), dtype=np.float32)
    go.jacobian_apply(z32, y32)
    assert rel(y32.download().astype(np.float64), o.jacobian_apply(z)) <= 2e-5


# Grids through the row kernel (default) and through the general kernel: aligned rows (nx a
# multiple of the workgroup chunk: 16 cells Q3, 32 Q2, 64 Q1) and rows whose last chunk is
# partial. SPD cases also run CG, whose iterations use the apply with the fused z.(Jz)
# partials.
ALIGNED = [
    ("q3_16x3x5_var_conv", dict(nx=16, ny=3, nz=5, degree=3, alpha=3.0, b=(0.5, -1.0, 0.25), c=0.2), 21),
    ("q3_16x2x6_react", dict(nx=16, ny=2, nz=6, degree=3, alpha=3.0, c=0.5), None),
    ("q3_16x4x5_var", dict(nx=16, ny=4, nz=5, degree=3, alpha=3.0), 24),
    ("q3_32x2x3", dict(nx=32, ny=2, nz=3, degree=3, alpha=3.0), None),
    ("q3_32x1x3_neumann", dict(nx=32, ny=1, nz=3, degree=3, alpha=3.0, bctype=(0, -1, 0, 0, -1, 0)), None),
    ("q3_16x3x2_aniso_iipg", dict(nx=16, ny=3, nz=2, degree=3, method=2, lx=2.0, ly=0.5, lz=1.0, alpha=4.0), None),
    ("q2_32x2x7", dict(nx=32, ny=2, nz=7, degree=2, alpha=3.0), None),
    ("q2_64x3x3_nipg_var", dict(nx=64, ny=3, nz=3, degree=2, method=0, alpha=2.0), 22),
    ("q2_32x5x9_off", dict(nx=32, ny=5, nz=9, degree=2, weights=1, alpha=3.0), None),
    ("q1_64x3x4_var_conv", dict(nx=64, ny=3, nz=4, degree=1, alpha=3.0, b=(0.0, 1.0, -1.0)), 23),
    ("q1_64x1x1_off", dict(nx=64, ny=1, nz=1, degree=1, weights=1, alpha=3.0), None),
    ("q1_64x2x3_react_outflow", dict(nx=64, ny=2, nz=3, degree=1, alpha=3.0, c=0.4, b=(1.0, 0.0, 0.0),
                                     bctype=(0, -2, 0, 0, -1, 0)), None),
    # partial last chunks. The row kernel takes them (nx > chunk) for constant kappa without
    # convection when at most 1/8 (Q3) or 1/16 (Q2) of the row's lanes are wasted (nz % 2 at Q3, nz % 4 or
    # nz % 8 at Q2 adds the z-walk); the other grids go to the general kernel in both runs
    ("ragged_q3_31x3x3", dict(nx=31, ny=3, nz=3, degree=3, alpha=3.0), None),
    ("ragged_q3_30x2x4_walk", dict(nx=30, ny=2, nz=4, degree=3, alpha=3.0), None),
    ("ragged_q3_46x2x2_walk_react", dict(nx=46, ny=2, nz=2, degree=3, alpha=3.0, c=0.3), None),
    ("ragged_q2_62x2x4_walk", dict(nx=62, ny=2, nz=4, degree=2, alpha=3.0), None),
    ("ragged_q2_62x2x3", dict(nx=62, ny=2, nz=3, degree=2, alpha=3.0), None),
    ("ragged_q2_94x2x8_walk8", dict(nx=94, ny=2, nz=8, degree=2, alpha=3.0), None),
    ("ragged_q3_5x3x4_react_general", dict(nx=5, ny=3, nz=4, degree=3, alpha=3.0, c=0.3), None),
    ("ragged_q3_17x2x3_var_conv_general", dict(nx=17, ny=2, nz=3, degree=3, alpha=3.0, b=(0.5, -1.0, 0.25)), 25),
    ("ragged_q2_31x3x3_var_general", dict(nx=31, ny=3, nz=3, degree=2, alpha=3.0), 26),
    ("ragged_q1_70x2x2_general", dict(nx=70, ny=2, nz=2, degree=1, alpha=3.0), None),
]
VARIANTS = [(f"{k}_{name}", kw, seed, k) for name, kw, seed in ALIGNED for k in ("aligned", "general")]


@pytest.mark.parametrize("name,kw,seed,variant", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_apply_kernel_variants(pkg, of, monkeypatch, name, kw, seed, variant):
    say(f"[gpu] apply kernel variant case={name}")
    if variant == "general":
        monkeypatch.setenv("PDELAB_GPU_APPLY", "general")
    else:
        monkeypatch.delenv("PDELAB_GPU_APPLY", raising=False)
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    rng = np.random.default_rng(43)
    z = rng.uniform(-1, 1, go.N)
    y0 = rng.uniform(-1, 1, go.N)
    zd, yd = go.vector(z), go.vector(y0)
    go.jacobian_apply(zd, yd)
    assert rel(yd.download(), o.jacobian_apply(z, y0.copy())) <= APPLY_RTOL
    yd.upload(y0)
    go.jacobian_apply_transpose(zd, yd)
    assert rel(yd.download(), y0 + o.jacobian_apply_transpose(z)) <= APPLY_RTOL
    z32 = go.vector(z.astype(np.float32), dtype=np.float32)
    y32 = go.vector(np.zeros(go.N, np.float32), dtype=np.float32)
    go.jacobian_apply(z32, y32)
    assert rel(y32.download().astype(np.float64), o.jacobian_apply(z)) <= 2e-5
    if kw.get("method", 1) == 1 and "b" not in kw:
        rhs = rng.normal(size=go.N)
        x, r = go.vector(), go.vector(rhs)
        res = go.solve(x, r, method=0, reduction=1e-10, maxit=5000)
        assert res["converged"]
        assert rel(o.jacobian_apply(x.download()), rhs) <= 1e-9


def test_apply_kernel_follows_set_kappa(pkg, of):
    """A partial-chunk grid with constant kappa takes the row kernel with the z-walk (12
    fused-dot partials); set_kappa makes the operator cell-wise, which sends the same grid
    to the general kernel (23 partials, more than the first solve's walk needed). Both
    CG solves (block Jacobi) follow the oracle. The iteration counts may differ by 1% on
    the lognormal operator, whose CG iteration count moves with rounding at the 1e-12 level
    (point Jacobi there: 864 against the oracle's 858 over ~860 iterations)."""
    cfg = pkg.make_config(nx=30, ny=3, nz=4, degree=3, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    o.set_threads(8)
    k = np.exp(np.random.default_rng(27).normal(size=go.ncells))
    for step in ("constant kappa", "cell-wise kappa"):
        say(f"[gpu] apply kernel choice 30x3x4 Q3, {step}")
        xd = go.vector()
        res = go.linear_problem_solve(xd, method=0, precond=2, reduction=1e-10)
        assert res["solver"]["converged"]
        x_ref, res_ref = o.linear_problem_solve(method=0, precond=2, reduction=1e-10)
        it, it_ref = int(res["solver"]["iterations"]), int(res_ref["solver"]["iterations"])
        assert abs(it - it_ref) <= max(2, 0.01 * it_ref), (step, it, it_ref)
        x = xd.download()
        assert rel(x, x_ref) <= 1e-7
        assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * res["first_defect"]
        go.set_kappa(k)
        o.set_kappa(k)


PROBLEMS = [
    ("exp_q2", dict(n=4, degree=2, alpha=3.0), None, 1),
    ("quadratic_q3_var", dict(nx=3, ny=2, nz=4, degree=3, alpha=3.0, quad=(0.3, 1.0, 2.0, -1.0, 1.0, 1.0, -2.0)), 9, 4),
    ("gauss_conv_q2", dict(n=4, degree=2, alpha=3.0, b=(1.0, 0.5, 0.25), kappa0=0.01,
                           gaussians=((0.3, 0.4, 0.5, 0.1, 1.0), (0.7, 0.2, 0.6, 0.15, -0.5)), g0=0.2), None, 3),
    ("constant_intorder_q1", dict(n=5, degree=1, alpha=2.0, f0=1.5, g0=-0.5, intorderadd=3), 8, 2),
    # Neumann flux j and outflow data o: the exact solution's fluxes (exp, quadratic) or the
    # constants j0 / o0 (convectiondiffusiondg.hh:776-783, 810-815)
    ("exp_neumann_outflow_q3", dict(nx=3, ny=4, nz=2, degree=3, alpha=3.0, b=(1.0, -0.5, 0.25),
                                    bctype=(0, -2, -2, -1, 0, -1)), 10, 1),
    ("quadratic_neumann_q2_var", dict(n=3, degree=2, alpha=3.0, quad=(0.3, 1.0, 2.0, -1.0, 1.0, 1.0, -2.0),
                                      bctype=(-1, 0, 0, -1, -1, 0)), 11, 4),
    ("constant_j0_o0_q2", dict(nx=4, ny=2, nz=3, degree=2, alpha=3.0, b=(0.0, 0.0, 2.0), f0=1.0, g0=0.5, j0=-1.5,
                               o0=0.75, bctype=(-1, 0, 0, -1, 0, -2)), None, 2),
    ("zero_problem_j0_q1", dict(n=4, degree=1, alpha=3.0, j0=2.0, bctype=(-1, -1, 0, 0, 0, 0)), None, 0),
]


@pytest.mark.parametrize("name,kw,seed,problem", PROBLEMS, ids=[p[0] for p in PROBLEMS])
def test_residual(pkg, of, name, kw, seed, problem):
    say(f"[gpu] residual parity case={name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed, problem=problem)
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, go.N)
    xd, rd = go.vector(x), go.vector()
    go.residual(xd, rd)
    r_ref = o.residual(x)
    assert rel(rd.download(), r_ref) <= APPLY_RTOL
    # residual at zero: the load vector alone
    rd.zero()
    go.residual(go.vector(), rd)
    assert rel(rd.download(), o.residual(np.zeros(go.N))) <= APPLY_RTOL


def test_tabulated_problem(pkg, of):
    say("[gpu] tabulated f/g")
    cfg = pkg.make_config(nx=3, ny=4, nz=2, degree=2, alpha=3.0, problem=pkg.PROBLEM_TABULATED)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    m = go.quadrature_points()
    assert m == o.quadrature_points()
    rng = np.random.default_rng(2)
    nb = 2 * (4 * 2 + 3 * 2 + 3 * 4)
    fq = rng.normal(size=go.ncells * m ** 3)
    gq = rng.normal(size=nb * m ** 2)
    go.set_tabulated(fq, gq)
    o.set_tabulated(fq, gq)
    rd = go.vector()
    x = rng.normal(size=go.N)
    go.residual(go.vector(x), rd)
    assert rel(rd.download(), o.residual(x)) <= APPLY_RTOL


def test_tabulated_neumann_outflow(pkg, of):
    """Tabulated boundary data on Neumann and outflow sides: the value of each face is j or o by
    the side's type."""
    say("[gpu] tabulated f / g / j / o")
    cfg = pkg.make_config(nx=3, ny=4, nz=2, degree=2, alpha=3.0, problem=pkg.PROBLEM_TABULATED, b=(0.5, 0.0, 0.0),
                          bctype=(0, -2, -1, 0, -1, 0))
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    m = go.quadrature_points()
    rng = np.random.default_rng(12)
    nb = 2 * (4 * 2 + 3 * 2 + 3 * 4)
    fq = rng.normal(size=go.ncells * m ** 3)
    gq = rng.normal(size=nb * m ** 2)
    go.set_tabulated(fq, gq)
    o.set_tabulated(fq, gq)
    rd = go.vector()
    x = rng.normal(size=go.N)
    go.residual(go.vector(x), rd)
    assert rel(rd.download(), o.residual(x)) <= APPLY_RTOL


def test_cell_source(pkg, of):
    say("[gpu] cell source")
    cfg = pkg.make_config(n=3, degree=2, alpha=3.0, problem=pkg.PROBLEM_CONSTANT, f0=1.0)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    s = np.random.default_rng(3).normal(size=27)
    go.set_cell_source(s)
    o.set_cell_source(s)
    rd = go.vector()
    go.residual(go.vector(), rd)
    assert rel(rd.download(), o.residual(np.zeros(go.N))) <= APPLY_RTOL


def test_vector_backend_deterministic(pkg, of):
    say("[gpu] dot/norm/axpy")
    cfg = pkg.make_config(n=20, degree=2)
    go = pkg.GridOperator(cfg)
    rng = np.random.default_rng(4)
    x, y = rng.normal(size=go.N), rng.normal(size=go.N)
    xd, yd = go.vector(x), go.vector(y)
    d1 = go.dot(xd, yd)
    d2 = go.dot(xd, yd)
    assert d1 == d2
    assert abs(d1 - x @ y) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)
    assert abs(go.norm(xd) - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    go.axpy(0.5, xd, yd)
    assert np.abs(yd.download() - (y + 0.5 * x)).max() <= 1e-15 * np.abs(y + 0.5 * x).max()


def test_apply_deterministic(pkg, of):
    say("[gpu] apply run-to-run bitwise")
    cfg = pkg.make_config(n=12, degree=3, alpha=3.0, kappa_per_cell=True)
    kappa = np.exp(np.random.default_rng(5).normal(size=12 ** 3))
    go = pkg.GridOperator(cfg, kappa)
    zd = go.vector(np.random.default_rng(6).uniform(-1, 1, go.N))
    a, b = go.vector(), go.vector()
    go.jacobian_apply(zd, a)
    go.jacobian_apply(zd, b)
    assert np.array_equal(a.download(), b.download())


@pytest.mark.parametrize("method", [0, 1], ids=["cg", "bicgstab"])
def test_linear_problem_solve_c1(pkg, of, method):
    """16^3 Q2 SIPG Poisson with the manufactured exp solution, alpha 3, CG/BiCGSTAB +
    point Jacobi to 1e-10: GPU solution vs oracle solution at the solver tolerance."""
    say(f"[gpu] linear problem 16^3 Q2 method={method}")
    cfg = pkg.make_config(n=16, degree=2, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=method, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=method, reduction=1e-10)
    x = xd.download()
    # both solve the same system to 1e-10 relative residual
    assert rel(x, x_ref) <= 1e-7
    assert abs(res["first_defect"] - res_ref["first_defect"]) <= 1e-12 * res_ref["first_defect"]
    # same Krylov recurrence; CG counts agree to 2, BiCGSTAB's are rounding-sensitive (15 %)
    it, it_ref = int(res["solver"]["iterations"]), int(res_ref["solver"]["iterations"])
    assert abs(it - it_ref) <= (2 if method == 0 else max(2, int(0.15 * it_ref)))
    assert o.l2_error_sq(x) < 1e-8
    # the residual of the returned solution, in FP64 on the oracle
    r = o.residual(x)
    assert np.linalg.norm(r) <= 1.01e-10 * res["first_defect"]


def test_solver_not_converged_and_errors(pkg, of):
    say("[gpu] error paths")
    cfg = pkg.make_config(n=4, degree=1, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    res = go.linear_problem_solve(go.vector(), reduction=1e-10, maxit=1)
    assert not res["solver"]["converged"]
    assert res["solver"]["iterations"] == 1
    with pytest.raises(pkg.PdelabError) as ei:
        go.jacobian_apply_nonlinear(go.vector(), go.vector(), go.vector())
    assert ei.value.status == 2
    # NaN defect -> solver abort (dune-istl SolverAbort)
    x = np.zeros(go.N)
    x[3] = np.nan
    r = go.vector(x)
    with pytest.raises(pkg.PdelabError) as ei:
        go.solve(go.vector(), r)
    assert ei.value.status == 5
    # outflow on inflow (convectiondiffusiondg.hh:802-806)
    cfg2 = pkg.make_config(n=2, degree=1, b=(-1.0, 0, 0), bctype=(0, -2, 0, 0, 0, 0))
    go2 = pkg.GridOperator(cfg2)
    with pytest.raises(pkg.PdelabError) as ei:
        go2.jacobian_apply(go2.vector(), go2.vector())
    assert ei.value.status == 6
    bad = pkg.make_config(n=2, degree=5)
    with pytest.raises(pkg.PdelabError):
        pkg.GridOperator(bad)


def test_mixed_precision_refinement(pkg, of):
    say("[gpu] mixed precision 12^3 Q3 lognormal kappa")
    n = 12
    cfg = pkg.make_config(n=n, degree=3, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0, g0=0.0)
    kappa = np.exp(np.random.default_rng(8).normal(size=n ** 3))
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    o.set_threads(8)
    xd = go.vector()
    res = go.solve_mixed(xd, inner_reduction=0.1, outer_reduction=1e-10, inner_maxit=20000)
    assert res["converged"], res
    x = xd.download()
    x_ref, _ = o.linear_problem_solve(reduction=1e-13, maxit=20000)
    assert rel(x, x_ref) <= 1e-8
    r = o.residual(x)
    assert np.linalg.norm(r) <= 1.01e-10 * res["first_defect"]


def test_mixed_precision_bicgstab_convection(pkg, of):
    """The 96^3 Q2 convection-diffusion bench problem at 8^3: FP32 BiCGSTAB inside FP64
    refinement to 1e-8 against the oracle's FP64 BiCGSTAB solve to 1e-12."""
    say("[gpu] mixed precision BiCGSTAB 8^3 Q2 convection")
    gauss = ((0.3, 0.4, 0.5, 0.1, 1.0), (0.7, 0.2, 0.6, 0.15, -0.5), (0.5, 0.8, 0.3, 0.08, 0.8))
    cfg = pkg.make_config(n=8, degree=2, alpha=3.0, kappa0=1e-2, b=(1.0, 0.5, 0.25), problem=pkg.PROBLEM_GAUSSIANS,
                          gaussians=gauss, g0=0.0)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    o.set_threads(8)
    xd = go.vector()
    res = go.solve_mixed(xd, method=pkg.SOLVER_BICGSTAB, inner_reduction=1e-3, outer_reduction=1e-8,
                         inner_maxit=20000)
    assert res["converged"], res
    x = xd.download()
    x_ref, _ = o.linear_problem_solve(method=1, reduction=1e-12, maxit=20000)
    r = o.residual(x)
    assert np.linalg.norm(r) <= 1.01e-8 * res["first_defect"]
    assert rel(x, x_ref) <= 1e-6


@pytest.mark.parametrize("variant", ["restarted_cg", "flexible_cg", "flexible_cg_block"])
def test_mixed_precision_variants_cg(pkg, of, variant):
    """The 64^3 lognormal-kappa problem at 12^3: restarted refinement (each FP32 CG to 1e-4) and
    flexible GMRES over FP32 CG, against the oracle's residual of the returned solution."""
    say(f"[gpu] mixed precision {variant} 12^3 Q3 lognormal kappa")
    n = 12
    cfg = pkg.make_config(n=n, degree=3, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0, g0=0.0)
    kappa = np.exp(np.random.default_rng(8).normal(size=n ** 3))
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    o.set_threads(8)
    flags = pkg.MIXED_RESTARTED if variant == "restarted_cg" else pkg.MIXED_FLEXIBLE
    precond = pkg.PREC_BLOCK_JACOBI if variant.endswith("block") else pkg.PREC_JACOBI
    xd = go.vector()
    res = go.solve_mixed(xd, precond=precond, variant=flags, outer_reduction=1e-10, inner_maxit=20000, outer_maxit=200)
    assert res["converged"], res
    x = xd.download()
    assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * res["first_defect"]
    x_ref, _ = o.linear_problem_solve(reduction=1e-13, maxit=20000)
    assert rel(x, x_ref) <= 1e-8


@pytest.mark.parametrize("precond", [1, 2], ids=["jacobi", "block_jacobi"])
def test_mixed_precision_flexible_bicgstab_convection(pkg, of, precond):
    """Flexible GMRES in FP64 over FP32 BiCGSTAB (to 1e-2 from zero) on the convection-diffusion
    bench problem at 8^3, to 1e-8: the oracle's residual of the solution and its distance to the
    oracle's FP64 BiCGSTAB solve to 1e-12."""
    say(f"[gpu] flexible mixed BiCGSTAB 8^3 Q2 convection precond={precond}")
    gauss = ((0.3, 0.4, 0.5, 0.1, 1.0), (0.7, 0.2, 0.6, 0.15, -0.5), (0.5, 0.8, 0.3, 0.08, 0.8))
    cfg = pkg.make_config(n=8, degree=2, alpha=3.0, kappa0=1e-2, b=(1.0, 0.5, 0.25), problem=pkg.PROBLEM_GAUSSIANS,
                          gaussians=gauss, g0=0.0)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    o.set_threads(8)
    xd = go.vector()
    res = go.solve_mixed(xd, method=pkg.SOLVER_BICGSTAB, precond=precond, variant=pkg.MIXED_FLEXIBLE,
                         outer_reduction=1e-8, inner_maxit=20000, outer_maxit=200)
    assert res["converged"], res
    x = xd.download()
    assert np.linalg.norm(o.residual(x)) <= 1.01e-8 * res["first_defect"]
    x_ref, _ = o.linear_problem_solve(method=1, reduction=1e-12, maxit=20000)
    assert rel(x, x_ref) <= 1e-6


def test_mixed_precision_cg_small_load(pkg, of):
    """Reliable-update CG with a load of 1e-9: the FP32 recursion runs on the residual at unit
    norm, so r.r does not underflow in FP32 and the solve converges to 1e-10 as with f = 1."""
    say("[gpu] mixed precision CG, f0 = 1e-9")
    n = 8
    cfg = pkg.make_config(n=n, degree=2, alpha=3.0, problem=pkg.PROBLEM_CONSTANT, f0=1e-9, g0=0.0)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    xd = go.vector()
    res = go.solve_mixed(xd, outer_reduction=1e-10, inner_maxit=20000)
    assert res["converged"], res
    assert np.linalg.norm(o.residual(xd.download())) <= 1.01e-10 * res["first_defect"]


def test_mixed_precision_variant_flags_rejected_together(pkg):
    go = pkg.GridOperator(pkg.make_config(n=2, degree=1))
    with pytest.raises(pkg.PdelabError) as e:
        go.solve_mixed(go.vector(), variant=pkg.MIXED_RESTARTED | pkg.MIXED_FLEXIBLE)
    assert e.value.status == 1
    go.close()


def test_adjoint_gradient_vs_oracle(pkg, of):
    say("[gpu] grad logkappa 6^3 Q2")
    n = 6
    cfg = pkg.make_config(n=n, degree=2, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0)
    rng = np.random.default_rng(9)
    kappa = np.exp(rng.normal(0, 0.5, n ** 3))
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    u = rng.normal(size=go.N)
    lam = rng.normal(size=go.N)
    g = go.cell_vector()
    go.grad_logkappa(go.vector(u), go.vector(lam), g)
    g_ref = o.grad_logkappa(u, lam)
    assert rel(g.download(), g_ref) <= 1e-12
    gs = go.cell_vector()
    go.grad_source(go.vector(lam), gs)
    assert rel(gs.download(), o.grad_source(lam)) <= 1e-12


GRAD_CASES = [
    ("weights_off_g0_q2", dict(n=4, degree=2, alpha=3.0, weights=1, problem=2, f0=1.0)),
    ("exp_g_q1", dict(n=4, degree=1, alpha=3.0, problem=1, bctype=(0, 0, -1, 0, 0, 0))),
    ("exp_g_weights_off_q3", dict(nx=3, ny=2, nz=3, degree=3, alpha=2.0, weights=1, problem=1,
                                  bctype=(0, -1, 0, 0, 0, 0))),
    ("tabulated_g_nipg_q2", dict(nx=3, ny=3, nz=2, degree=2, alpha=3.0, method=0, problem=5)),
]


@pytest.mark.parametrize("name,kw", GRAD_CASES, ids=[c[0] for c in GRAD_CASES])
def test_grad_logkappa_general_vs_oracle(pkg, of, name, kw):
    """The kappa gradient with weightsOff and with Dirichlet data g != 0 (the derivative of the
    residual's data terms included) against the oracle's dual-number derivative."""
    say(f"[gpu] grad logkappa case={name}")
    cfg = pkg.make_config(kappa_per_cell=True, **kw)
    rng = np.random.default_rng(21)
    kappa = np.exp(rng.normal(0, 0.5, pkg.ncells(cfg)))
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    if kw.get("problem") == 5:
        m = go.quadrature_points()
        nb = 2 * (cfg.nx * cfg.ny + cfg.nx * cfg.nz + cfg.ny * cfg.nz)
        fq, gq = rng.normal(size=go.ncells * m ** 3), rng.normal(size=nb * m ** 2)
        go.set_tabulated(fq, gq)
        o.set_tabulated(fq, gq)
    u = rng.normal(size=go.N)
    lam = rng.normal(size=go.N)
    g = go.cell_vector()
    go.grad_logkappa(go.vector(u), go.vector(lam), g)
    assert rel(g.download(), o.grad_logkappa(u, lam)) <= 1e-12


@pytest.mark.parametrize("method", [0, 1], ids=["cg", "bicgstab"])
def test_transpose_solve(pkg, of, method):
    say(f"[gpu] transpose solve method={method}")
    cfg = pkg.make_config(n=5, degree=2, alpha=3.0, b=(1.0, 0.5, 0.25), kappa0=0.1 if method else 1.0)
    if method == 0:
        cfg.b[0] = cfg.b[1] = cfg.b[2] = 0.0
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    rhs = np.random.default_rng(10).normal(size=go.N)
    z, r = go.vector(), go.vector(rhs)
    res = go.solve(z, r, method=method, reduction=1e-12, transpose=True)
    assert res["converged"]
    zt = z.download()
    back = o.jacobian_apply_transpose(zt)
    assert rel(back, rhs) <= 2e-12


def test_large_apply_properties(pkg, of):
    """64^3 Q3 (16.8M DoF): linearity and symmetry x^T J y == y^T J x of the SIPG
    operator (the oracle comparison at this size is in test_gpu_fullsize.py)."""
    say("[gpu] 64^3 Q3 properties")
    cfg = pkg.make_config(n=64, degree=3, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, go.N)
    y = rng.uniform(-1, 1, go.N)
    xd, yd = go.vector(x), go.vector(y)
    jx, jy = go.vector(), go.vector()
    go.jacobian_apply(xd, jx)
    go.jacobian_apply(yd, jy)
    a = go.dot(yd, jx)
    b = go.dot(xd, jy)
    assert abs(a - b) <= 1e-11 * (abs(a) + abs(b))
    s = go.vector(x + 2.0 * y)
    js = go.vector()
    go.jacobian_apply(s, js)
    lin = js.download() - (jx.download() + 2.0 * jy.download())
    assert np.linalg.norm(lin) <= 1e-13 * np.linalg.norm(js.download())


def test_large_convection_properties(pkg):
    """96^3 Q2 convection-diffusion (23.9M DoF, the FP32-inner BiCGSTAB workload): duality
    <y, J x> == <J^T y, x>, linearity, and the FP32 apply against the FP64 one."""
    say("[gpu] 96^3 Q2 convection properties")
    cfg = pkg.make_config(n=96, degree=2, alpha=3.0, kappa0=1e-2, b=(1.0, 0.5, 0.25))
    go = pkg.GridOperator(cfg)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, go.N)
    y = rng.uniform(-1, 1, go.N)
    xd, yd = go.vector(x), go.vector(y)
    jx, jty = go.vector(), go.vector()
    go.jacobian_apply(xd, jx)
    go.jacobian_apply_transpose(yd, jty)
    a = go.dot(yd, jx)
    b = go.dot(jty, xd)
    assert abs(a - b) <= 1e-11 * (abs(a) + abs(b))
    s = go.vector(x - 0.5 * y)
    js, jy = go.vector(), go.vector()
    go.jacobian_apply(s, js)
    go.jacobian_apply(yd, jy)
    jxh = jx.download()
    lin = js.download() - (jxh - 0.5 * jy.download())
    assert np.linalg.norm(lin) <= 1e-13 * np.linalg.norm(jxh)
    x32, j32 = go.vector(x.astype(np.float32), dtype=np.float32), go.vector(dtype=np.float32)
    go.jacobian_apply(x32, j32)
    assert rel(j32.download().astype(np.float64), jxh) <= 1e-5
    go.close()


def test_large_lognormal_properties(pkg):
    """64^3 Q3 with cell-wise lognormal kappa (the mixed-precision workload): symmetry and
    linearity of the SIPG operator at full size."""
    say("[gpu] 64^3 Q3 lognormal kappa properties")
    n = 64
    cfg = pkg.make_config(n=n, degree=3, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0, g0=0.0)
    go = pkg.GridOperator(cfg, np.exp(np.random.default_rng(13).normal(size=n ** 3)))
    rng = np.random.default_rng(14)
    x = rng.uniform(-1, 1, go.N)
    y = rng.uniform(-1, 1, go.N)
    xd, yd = go.vector(x), go.vector(y)
    jx, jy = go.vector(), go.vector()
    go.jacobian_apply(xd, jx)
    go.jacobian_apply(yd, jy)
    a = go.dot(yd, jx)
    b = go.dot(xd, jy)
    assert abs(a - b) <= 1e-11 * (abs(a) + abs(b))
    s = go.vector(2.0 * x + y)
    js = go.vector()
    go.jacobian_apply(s, js)
    lin = js.download() - (2.0 * jx.download() + jy.download())
    assert np.linalg.norm(lin) <= 1e-13 * np.linalg.norm(js.download())
    go.close()


def test_vector_arguments_checked_before_the_call(pkg):
    """The C-ABI takes raw pointers; the binding rejects a vector of the wrong dtype or length
    on the host, before the library could read or write past the buffer."""
    say("[gpu] binding argument checks")
    go = pkg.GridOperator(pkg.make_config(nx=16, ny=3, nz=4, degree=3, alpha=3.0))
    z32, y32 = go.vector(dtype=np.float32), go.vector(dtype=np.float32)
    with pytest.raises(TypeError):
        go.operator_apply(z32, y32)          # FP64-only entry point
    with pytest.raises(TypeError):
        go.jacobian_apply(go.vector(), y32)  # mixed dtypes
    short = pkg.DeviceVector(go, go.N - 1)
    with pytest.raises(ValueError):
        go.residual(short, go.vector())
    with pytest.raises(ValueError):
        go.grad_source(go.vector(), go.vector())  # needs one value per cell
    go.jacobian_apply(z32, y32)              # the FP32 entry point accepts them
    go.synchronize()
    go.close()
