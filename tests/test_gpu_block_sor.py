"""GPU parity of the block-SOR preconditioner (through the C-ABI).

* pdelab_gpu_block_sor_apply (one forward sweep, one launch per hyperplane x + y + z), from
  v = 0 and in place from a given v, against the oracle's sequential sweep over the cells in
  index order, on every
  operator variant of the apply parity cases (constant and cell-wise kappa, SIPG / NIPG /
  IIPG, weightsOff, convection with reaction, outflow and Neumann sides, a single cell, a
  line of cells) and two relaxation factors: within 1e-10 relative (the dense inverse
  blocks reach condition numbers of 1e4 at Q3, as for block Jacobi);
* BiCGSTAB with block SOR against the oracle's solve with the same preconditioner:
  iteration count within 15 % (BiCGSTAB's count moves with rounding), solution at the
  solver tolerance, the oracle's residual of the solution;
* mixed precision (FP32 BiCGSTAB with the FP32 sweep inside FP64 refinement): the oracle's
  residual at the outer tolerance;
* the rejections: CG, the transposed operator, omega outside (0, 2), aliased vectors.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES, make, rel, say

pytestmark = pytest.mark.gpu

SOR = 3  # PDELAB_PREC_BLOCK_SOR
BJ = 2


@pytest.mark.parametrize("name,kw,seed", CASES, ids=[c[0] for c in CASES])
def test_block_sor_apply_vs_oracle(pkg, of, name, kw, seed):
    say(f"[gpu] block SOR apply {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    rng = np.random.default_rng(23)
    d = rng.uniform(-1, 1, go.N)
    v0 = rng.normal(size=go.N)
    dd, v = go.vector(d), go.vector()
    for omega in (1.0, 1.35):
        go.set_sor_relaxation(omega)
        o.set_sor_omega(omega)
        v.zero()
        go.block_sor_apply(dd, v)  # on v = 0: the preconditioner the solvers apply
        e = rel(v.download(), o.block_sor_apply(d))
        assert e <= 1e-10, (omega, e)
        v.upload(v0)
        go.block_sor_apply(dd, v)  # one SOR step from v0, in place
        e = rel(v.download(), o.block_sor_apply(d, v0))
        assert e <= 1e-10, (omega, e)
    go.close()


@pytest.mark.parametrize("name,kw,seed", [
    ("q2_conv_8", dict(n=8, degree=2, alpha=3.0, kappa0=0.01, b=(1.0, 0.5, 0.25), problem=2, f0=1.0), None),
    ("q3_var_6", dict(n=6, degree=3, alpha=3.0, problem=1), 21),
    ("q1_conv_var_9x7x5", dict(nx=9, ny=7, nz=5, degree=1, alpha=3.0, b=(-0.5, 1.0, 0.3), problem=2, f0=1.0), 4),
], ids=["q2_conv_8", "q3_var_6", "q1_conv_var_9x7x5"])
def test_block_sor_bicgstab_vs_oracle(pkg, of, name, kw, seed):
    say(f"[gpu] block SOR BiCGSTAB {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=1, precond=SOR, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=1, precond=SOR, reduction=1e-10)
    assert res_ref["solver"]["converged"]
    it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
    say(f"  iterations {it} (oracle {it_ref})")
    assert abs(it - it_ref) <= max(2, 0.15 * it_ref)
    x = xd.download()
    assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * res["first_defect"]
    assert rel(x, x_ref) <= 1e-7
    # the same handle then runs block Jacobi (its own data slot) and SOR again
    xb = go.vector()
    rb = go.linear_problem_solve(xb, method=1, precond=BJ, reduction=1e-10)
    assert rb["solver"]["converged"]
    xs = go.vector()
    rs = go.linear_problem_solve(xs, method=1, precond=SOR, reduction=1e-10)
    assert abs(rs["solver"]["iterations"] - it) <= max(2, 0.15 * it)
    go.close()


def test_block_sor_mixed_bicgstab(pkg, of):
    say("[gpu] block SOR mixed-precision BiCGSTAB 8^3 Q2 convection")
    kw = dict(n=8, degree=2, alpha=3.0, kappa0=0.01, b=(1.0, 0.5, 0.25), problem=2, f0=1.0)
    cfg, kappa, go, o = make(pkg, of, kw, None)
    xm = go.vector()
    rm = go.solve_mixed(xm, method=1, precond=SOR, inner_reduction=1e-4, outer_reduction=1e-10)
    assert rm["converged"], rm
    r0 = np.linalg.norm(o.residual(np.zeros(go.N)))
    assert np.linalg.norm(o.residual(xm.download())) <= 1.01e-10 * r0
    xf = go.vector()
    rf = go.solve_mixed(xf, method=1 | pkg.MIXED_FLEXIBLE, precond=SOR, outer_reduction=1e-10)
    assert rf["converged"], rf
    assert np.linalg.norm(o.residual(xf.download())) <= 1.01e-10 * r0
    go.close()


def test_block_sor_rejections(pkg):
    say("[gpu] block SOR rejections")
    cfg = pkg.make_config(n=3, degree=1, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    with pytest.raises(RuntimeError, match="BiCGSTAB only"):
        go.linear_problem_solve(go.vector(), method=0, precond=SOR)
    with pytest.raises(RuntimeError, match="BiCGSTAB only"):
        go.solve_mixed(go.vector(), method=0, precond=SOR)
    with pytest.raises(RuntimeError, match="transposed"):
        go.solve(go.vector(), go.vector(np.ones(go.N)), method=1, precond=SOR, transpose=True)
    for bad in (0.0, 2.0, -1.0, float("nan")):
        with pytest.raises(RuntimeError):
            go.set_sor_relaxation(bad)
    d = go.vector(np.ones(go.N))
    with pytest.raises(RuntimeError):
        go.block_sor_apply(d, d)
    go.close()
