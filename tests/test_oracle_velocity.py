"""Pins the oracle's velocity field b(x) (tabulated at the quadrature points, the values
param.b(cell, x) returns where convectiondiffusiondg.hh evaluates it: :178 volume, :426
skeleton, :797 boundary):

  * a constant field gives the operator, residual and load vector of the constant b;
  * consistency: with a divergence-free field linear in x, the interpolant of a Q_k
    polynomial has zero residual (the manufactured f = -kappa lap u + b.grad u + c u and
    j = (b.n) u - kappa grad u.n use the field at the points), with Dirichlet, Neumann and
    outflow sides, SIPG and NIPG, Q1-Q3, also with the rule of order 2k + 2. Every integral is
    exact there (the integrands have degree <= 2k + 1 per direction), so the residual is
    rounding; a field read at the wrong point (swapped tangential directions of the face
    points, the volume components permuted, the face of the other side) leaves an order-one
    residual, which the negative controls check;
  * structure with a random field: assembled matrix x == jacobian_apply(x), the point
    diagonal, the transpose, the two-sided traversal;
  * an outflow side that sees inflow is refused.
"""
import numpy as np
import pytest


def div_free(x, y, z):
    """b = (0.5 - (y - 1/2) + 0.3 z, 0.25 + (x - 1/2) - 0.2 z, 0.2 + 0.5 (x - 1/2) - 0.3 (y - 1/2)):
    linear, divergence-free, b_x >= 0 on the unit cube."""
    return (0.5 - (y - 0.5) + 0.3 * z, 0.25 + (x - 0.5) - 0.2 * z, 0.2 + 0.5 * (x - 0.5) - 0.3 * (y - 0.5))


def test_constant_field_equals_constant_b(of):
    cm = of.cfgmod
    b = (1.0, 0.5, -0.25)
    kw = dict(nx=3, ny=2, nz=2, degree=2, alpha=3.0, c=0.4, kappa0=0.3, problem=cm.PROBLEM_EXP,
              bctype=(0, -2, -1, 0, 0, -1))
    oc = of.Oracle(cm.make_config(b=b, **kw))
    cfg = cm.make_config(**kw)
    ov = of.Oracle(cfg)
    bv, bf = cm.tabulate_velocity(cfg, lambda x, y, z: b)
    ov.set_velocity_field(bv, bf)
    A, Av = oc.dense_matrix(), ov.dense_matrix()
    assert np.abs(A - Av).max() <= 1e-14 * np.abs(A).max()
    x = np.random.default_rng(3).normal(size=oc.N)
    r = oc.residual(x)
    assert np.abs(ov.residual(x) - r).max() <= 1e-14 * np.abs(r).max()
    # cleared: back to b = 0
    ov.set_velocity_field(None, None)
    assert np.abs(ov.dense_matrix() - of.Oracle(cfg).dense_matrix()).max() == 0.0


REPRO = [(1, 1, 0), (2, 1, 0), (3, 1, 0), (2, 0, 0), (2, 1, 2)]


@pytest.mark.parametrize("k,method,intorderadd", REPRO,
                         ids=["q1_sipg", "q2_sipg", "q3_sipg", "q2_nipg", "q2_sipg_intorderadd2"])
def test_polynomial_reproduction_with_field(of, k, method, intorderadd):
    cm = of.cfgmod
    q = (0.3, 1.0, 2.0, -1.0, 0.0, 0.0, 0.0) if k == 1 else (0.3, 1.0, 2.0, -1.0, 1.0, 1.5, -2.0)
    # +x outflow (b_x >= 0), -y and +z Neumann, the rest Dirichlet
    cfg = cm.make_config(nx=3, ny=2, nz=3, degree=k, method=method, alpha=4.0, kappa0=0.7, c=0.3,
                         intorderadd=intorderadd, problem=cm.PROBLEM_QUADRATIC, quad=q, bctype=(0, -2, -1, 0, 0, -1))
    o = of.Oracle(cfg)
    bv, bf = cm.tabulate_velocity(cfg, div_free)
    o.set_velocity_field(bv, bf)
    u = o.interpolate()
    assert np.abs(o.residual(u)).max() < 1e-12
    # negative controls: the field read at the wrong points
    m = cm.quadrature_points(cfg)
    o.set_velocity_field(bv, bf.reshape(-1, m, m).transpose(0, 2, 1).ravel())
    assert np.abs(o.residual(u)).max() > 1e-3
    o.set_velocity_field(bv.reshape(-1, 3)[:, [1, 2, 0]].ravel(), bf)
    assert np.abs(o.residual(u)).max() > 1e-3


def test_solution_of_field_problem(of):
    """BiCGSTAB recovers the interpolant (the discrete solution) with the field."""
    cm = of.cfgmod
    q = (0.3, 1.0, 2.0, -1.0, 1.0, 1.5, -2.0)
    cfg = cm.make_config(nx=3, ny=3, nz=2, degree=2, alpha=4.0, kappa0=0.5, problem=cm.PROBLEM_QUADRATIC, quad=q)
    o = of.Oracle(cfg)
    o.set_velocity_field(*cm.tabulate_velocity(cfg, div_free))
    x, res = o.linear_problem_solve(method=1, reduction=1e-13, maxit=5000)
    assert res["solver"]["converged"]
    assert np.abs(x - o.interpolate()).max() < 1e-8


def random_field(cfg, seed):
    cm_m = (2 * cfg.degree + cfg.intorderadd) // 2 + 1
    rng = np.random.default_rng(seed)
    nc = cfg.nx * cfg.ny * cfg.nz
    nf = (cfg.nx + 1) * cfg.ny * cfg.nz + cfg.nx * (cfg.ny + 1) * cfg.nz + cfg.nx * cfg.ny * (cfg.nz + 1)
    return rng.normal(size=nc * cm_m ** 3 * 3), rng.normal(size=nf * cm_m ** 2)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_structure_with_random_field(of, k):
    cm = of.cfgmod
    cfg = cm.make_config(nx=3, ny=2, nz=2, degree=k, alpha=3.0, kappa0=0.2, c=0.1, bctype=(0, 0, -1, 0, 0, -3))
    o = of.Oracle(cfg)
    o.set_velocity_field(*random_field(cfg, 40 + k))
    A = o.dense_matrix()
    x = np.random.default_rng(41).uniform(-1, 1, o.N)
    y = o.jacobian_apply(x)
    s = np.abs(y).max()
    assert np.abs(y - A @ x).max() <= 1e-13 * s
    assert np.abs(o.point_diagonal() - np.diag(A)).max() <= 1e-13 * np.abs(A).max()
    assert np.abs(o.jacobian_apply_transpose(x) - A.T @ x).max() <= 1e-13 * s
    assert np.abs(o.jacobian_apply_twosided(x, 2) - y).max() <= 1e-13 * s
    assert np.abs(A - A.T).max() > 1e-3  # the convection terms are not symmetric


def test_outflow_on_inflow_refused(of):
    cm = of.cfgmod
    cfg = cm.make_config(nx=2, ny=2, nz=2, degree=1, alpha=3.0, bctype=(-2, 0, 0, 0, 0, 0))
    o = of.Oracle(cfg)
    o.set_velocity_field(*cm.tabulate_velocity(cfg, div_free))  # b_x > 0: inflow on the -x side
    with pytest.raises(RuntimeError, match="Outflow"):
        o.residual(np.zeros(o.N))
