"""GPU parity of the instationary path (OneStepGridOperator over the operator and the DG mass,
OneStepMethod::apply):

* pdelab_gpu_onestep_jacobian_apply against the oracle's b J z + a M z, which evaluates the
  operator and the ScalarL2 mass separately as the reference's OneStepGridOperator does (the
  GPU folds the mass into the reaction term of the apply kernels): within 1e-12 relative;
* pdelab_gpu_onestep_step for implicit Euler, Crank-Nicolson and Alexander2, with CG/point
  Jacobi, CG/block Jacobi and BiCGSTAB (convection), against the oracle's step: solutions
  within 1e-8 relative after two steps (both solve each stage to 1e-12), stage iteration
  totals within 15 %.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES, make, rel, say

pytestmark = pytest.mark.gpu

SOME = [c for c in CASES if c[0] in ("sipg_q2_5x4x3", "sipg_q3_var", "nipg_q2_var", "conv_react_q2",
                                     "row_q3_16x2x4", "row_q2_32x3x2_var", "outflow_q2")]


@pytest.mark.parametrize("name,kw,seed", SOME, ids=[c[0] for c in SOME])
def test_onestep_jacobian_apply(pkg, of, name, kw, seed):
    say(f"[gpu] onestep apply {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    rng = np.random.default_rng(41)
    z, y0 = rng.uniform(-1, 1, go.N), rng.uniform(-1, 1, go.N)
    for a_mass, b_dt in ((1.0, 0.05), (0.0, 2.0), (3.0, 0.0), (1.0, 1e-4)):
        yd = go.vector(y0)
        go.onestep_jacobian_apply(a_mass, b_dt, go.vector(z), yd)
        want = o.onestep_jacobian_apply(a_mass, b_dt, z, y0.copy())
        assert rel(yd.download(), want) <= 1e-12, (a_mass, b_dt)


STEP_CASES = [
    ("ie_cg_jacobi", dict(n=5, degree=2, alpha=3.0, problem=1), None, 0, 0, 1),
    ("cn_cg_block_jacobi", dict(n=4, degree=3, alpha=3.0, problem=1), None, 1, 0, 2),
    ("alex2_cg_jacobi_var", dict(n=5, degree=2, alpha=3.0, problem=2, f0=1.0), 3, 2, 0, 1),
    ("alex2_cg_block_jacobi_var", dict(n=4, degree=3, alpha=3.0, problem=2, f0=1.0), 4, 2, 0, 2),
    ("alex2_bicgstab_conv", dict(n=5, degree=2, alpha=3.0, kappa0=0.05, b=(1.0, 0.5, 0.25), problem=2, f0=1.0),
     None, 2, 1, 1),
]


@pytest.mark.parametrize("name,kw,seed,scheme,method,precond", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_onestep_step(pkg, of, name, kw, seed, scheme, method, precond):
    say(f"[gpu] onestep step {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    o.set_threads(8)
    x0 = np.random.default_rng(42).normal(size=go.N) * 0.1
    dt = 0.05
    xo, xn = go.vector(x0), go.vector(x0)
    x_ref = x0.copy()
    its = its_ref = 0
    for _ in range(2):
        res = go.onestep_step(xo, xn, scheme=scheme, dt=dt, method=method, precond=precond, reduction=1e-12)
        assert res["converged"] and res["stages"] == (2 if scheme == 2 else 1)
        xo.copy_from(xn)
        x_ref, rr = o.onestep_step(x_ref, scheme, dt, xnew=x_ref, method=method, precond=precond, reduction=1e-12)
        assert rr["converged"]
        its += res["linear_solver_iterations"]
        its_ref += rr["linear_solver_iterations"]
    assert rel(xn.download(), x_ref) <= 1e-8
    assert abs(its - its_ref) <= max(2, 0.15 * its_ref)
