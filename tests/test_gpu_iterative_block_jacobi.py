"""GPU parity of the iterative block Jacobi preconditioner (through the C-ABI).

* pdelab_gpu_iterative_block_jacobi_apply (v += W^{-1} d; one wave per cell running dune-istl's BiCGSTAB on
  the cell's diagonal block, point-Jacobi preconditioned) against the oracle's per-cell solves
  on every operator variant of the apply parity cases: within 1e-4 relative (a block whose
  defect ends next to the 1e-5 threshold may stop one half step apart under different rounding,
  which moves its solution by about the threshold), with the block solvers' iteration counts
  summed over the cells within 2 %; driven to a 1e-11 reduction, within 1e-7 of the oracle;
* BiCGSTAB with the preconditioner against the oracle's solve with it: iteration count within
  15 %, solution at the solver tolerance, the oracle's residual of the solution; the
  mixed-precision BiCGSTAB (FP32 block solves) to the outer tolerance;
* the rejections: CG, the transposed operator, invalid options, aliased vectors.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES, make, rel, say

pytestmark = pytest.mark.gpu

IBJ = 4  # PDELAB_PREC_ITERATIVE_BLOCK_JACOBI


@pytest.mark.parametrize("name,kw,seed", CASES, ids=[c[0] for c in CASES])
def test_iterative_block_jacobi_apply_vs_oracle(pkg, of, name, kw, seed):
    say(f"[gpu] iterative block Jacobi apply {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    d = np.random.default_rng(29).uniform(-1, 1, go.N)
    dd, v = go.vector(d), go.vector(np.ones(go.N))
    its = go.iterative_block_jacobi_apply(dd, v)  # accumulates, as GridOperatorPreconditioner::apply
    want, its_ref = o.iterative_block_jacobi_apply(d)
    e = rel(v.download() - 1.0, want)
    say(f"  rel err {e:.2e}, block iterations {its} (oracle {its_ref})")
    assert e <= 1e-4
    assert abs(its - its_ref) <= max(2, its_ref // 50)
    go.set_block_solver_options(1e-11, 500, True, 1.0)
    o.set_block_solver(1e-11, 500, True, 1.0)
    v.zero()
    go.iterative_block_jacobi_apply(dd, v)
    want, _ = o.iterative_block_jacobi_apply(d)
    assert rel(v.download(), want) <= 1e-7
    go.close()


@pytest.mark.parametrize("name,kw,seed", [
    ("q2_conv_6", dict(n=6, degree=2, alpha=3.0, kappa0=0.01, b=(1.0, 0.5, 0.25), problem=2, f0=1.0), None),
    ("q3_var_5", dict(n=5, degree=3, alpha=3.0, problem=1), 21),
], ids=["q2_conv_6", "q3_var_5"])
def test_iterative_block_jacobi_bicgstab_vs_oracle(pkg, of, name, kw, seed):
    say(f"[gpu] iterative block Jacobi BiCGSTAB {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=1, precond=IBJ, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=1, precond=IBJ, reduction=1e-10)
    assert res_ref["solver"]["converged"]
    it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
    say(f"  iterations {it} (oracle {it_ref})")
    assert abs(it - it_ref) <= max(2, 0.15 * it_ref)
    x = xd.download()
    assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * res["first_defect"]
    assert rel(x, x_ref) <= 1e-7
    xm = go.vector()
    rm = go.solve_mixed(xm, method=1, precond=IBJ, inner_reduction=1e-4, outer_reduction=1e-10)
    assert rm["converged"], rm
    assert np.linalg.norm(o.residual(xm.download())) <= 1.01e-10 * res["first_defect"]
    go.close()


def test_iterative_block_jacobi_rejections(pkg):
    say("[gpu] iterative block Jacobi rejections")
    cfg = pkg.make_config(n=3, degree=1, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    with pytest.raises(RuntimeError, match="BiCGSTAB only"):
        go.linear_problem_solve(go.vector(), method=0, precond=IBJ)
    with pytest.raises(RuntimeError, match="BiCGSTAB only"):
        go.solve_mixed(go.vector(), method=0, precond=IBJ)
    with pytest.raises(RuntimeError, match="transposed"):
        go.solve(go.vector(), go.vector(np.ones(go.N)), method=1, precond=IBJ, transpose=True)
    for bad in ((0.0, 10, 1, 1.0), (1e-5, 0, 1, 1.0), (1e-5, 10, 1, 0.0), (float("nan"), 10, 1, 1.0)):
        with pytest.raises(RuntimeError):
            go.set_block_solver_options(*bad)
    d = go.vector(np.ones(go.N))
    with pytest.raises(RuntimeError):
        go.iterative_block_jacobi_apply(d, d)
    go.close()
