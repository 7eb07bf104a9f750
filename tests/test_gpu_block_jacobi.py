"""GPU parity of the per-cell block Jacobi preconditioner (through the C-ABI).

* pdelab_gpu_block_diagonal against the oracle's BlockDiagonalLocalOperatorWrapper blocks
  (1e-12 relative), for J and J^T, on every operator variant of the apply parity cases;
* pdelab_gpu_block_jacobi_apply against numpy.linalg.solve with the oracle's blocks:
  FP64 within 1e-10 relative, FP32 within 1e-4 relative (the blocks' condition numbers
  reach 1e4 at Q3, so the inverse loses up to four digits of the arithmetic);
* CG and BiCGSTAB with block Jacobi against the oracle's solves with the same
  preconditioner: iteration counts (CG within 2, BiCGSTAB within 15 %) and solutions at
  the solver tolerance; the transposed solve; mixed precision CG with block Jacobi.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES, make, rel, say

pytestmark = pytest.mark.gpu

BJ = 2  # PDELAB_PREC_BLOCK_JACOBI


def _solve_blocks(blocks, z, nloc):
    return np.linalg.solve(blocks, z.reshape(-1, nloc, 1))[..., 0].ravel()


@pytest.mark.parametrize("name,kw,seed", CASES, ids=[c[0] for c in CASES])
def test_block_diagonal_and_apply(pkg, of, name, kw, seed):
    say(f"[gpu] block jacobi {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    nloc = (cfg.degree + 1) ** 3
    ref = o.block_diagonal()
    bd = pkg.DeviceVector(go, go.ncells * nloc * nloc)
    go.block_diagonal(bd)
    got = bd.download().reshape(ref.shape)
    assert rel(got, ref) <= 1e-12
    go.block_diagonal(bd, transpose=True)
    assert rel(bd.download().reshape(ref.shape), np.transpose(ref, (0, 2, 1))) <= 1e-12
    rng = np.random.default_rng(17)
    z = rng.uniform(-1, 1, go.N)
    want = _solve_blocks(ref, z, nloc)
    y0 = rng.normal(size=go.N)
    y = go.vector(y0)
    go.block_jacobi_apply(go.vector(z), y)
    assert rel(y.download() - y0, want) <= 1e-10
    y32 = go.vector(np.zeros(go.N, dtype=np.float32), dtype=np.float32)
    go.block_jacobi_apply(go.vector(z.astype(np.float32), dtype=np.float32), y32)
    assert rel(y32.download().astype(np.float64), want) <= 1e-4


@pytest.mark.parametrize("seed", [None, 21], ids=["const", "lognormal"])
def test_block_jacobi_cg(pkg, of, seed):
    say(f"[gpu] block jacobi CG 6^3 Q3 seed={seed}")
    kw = dict(n=6, degree=3, alpha=3.0, problem=pkg.PROBLEM_EXP)
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=0, precond=BJ, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=0, precond=BJ, reduction=1e-10)
    it, it_ref = int(res["solver"]["iterations"]), int(res_ref["solver"]["iterations"])
    assert abs(it - it_ref) <= 2
    x = xd.download()
    assert rel(x, x_ref) <= 1e-7
    r = o.residual(x)
    assert np.linalg.norm(r) <= 1.01e-10 * res["first_defect"]
    # point Jacobi needs clearly more iterations on the same problem
    xp = go.vector()
    rp = go.linear_problem_solve(xp, method=0, precond=pkg.PREC_JACOBI, reduction=1e-10)
    assert it < 0.85 * int(rp["solver"]["iterations"])


# nx ny a multiple of the cells per wave of the block-Jacobi kernel (4 Q3, 7 Q2, 16 Q1)
# takes the kernel whose z entries are wave-uniform; the other grids the general one
GRIDS = [("q3_5x5x3_general", (5, 5, 3), 3), ("q3_4x3x4_uniform_z", (4, 3, 4), 3),
         ("q2_7x3x2_uniform_z", (7, 3, 2), 2), ("q2_4x4x3_general", (4, 4, 3), 2),
         ("q1_8x2x3_uniform_z", (8, 2, 3), 1), ("q1_5x3x4_general", (5, 3, 4), 1),
         # aligned rows whose apply walks 2 (Q3) or 4 or 8 (Q2) z-chunks per workgroup
         ("q3_16x3x4_walk", (16, 3, 4), 3), ("q2_32x2x4_walk", (32, 2, 4), 2),
         # and rows whose last chunk is partial (row kernel), with and without the walk
         ("q3_30x3x4_walk_ragged", (30, 3, 4), 3), ("q2_62x2x3_ragged", (62, 2, 3), 2),
         ("q2_32x2x8_walk8", (32, 2, 8), 2)]


@pytest.mark.parametrize("name,dims,k", GRIDS, ids=[g[0] for g in GRIDS])
def test_block_jacobi_cg_grids(pkg, of, name, dims, k):
    say(f"[gpu] block jacobi CG {name}")
    kw = dict(nx=dims[0], ny=dims[1], nz=dims[2], degree=k, alpha=3.0, problem=pkg.PROBLEM_EXP)
    cfg, kappa, go, o = make(pkg, of, kw, None)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=0, precond=BJ, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=0, precond=BJ, reduction=1e-10)
    assert abs(int(res["solver"]["iterations"]) - int(res_ref["solver"]["iterations"])) <= 2
    assert rel(xd.download(), x_ref) <= 1e-7
    # the FP32 kernels of the same grid, inside the mixed-precision refinement
    xm = go.vector()
    rm = go.solve_mixed(xm, precond=BJ, inner_reduction=0.1, outer_reduction=1e-10)
    assert rm["converged"]
    assert rel(xm.download(), x_ref) <= 1e-7


def test_block_jacobi_bicgstab_convection(pkg, of):
    say("[gpu] block jacobi BiCGSTAB 6^3 Q2 convection")
    kw = dict(n=6, degree=2, alpha=3.0, kappa0=0.01, b=(1.0, 0.5, 0.25), problem=pkg.PROBLEM_CONSTANT, f0=1.0)
    cfg, kappa, go, o = make(pkg, of, kw, None)
    o.set_threads(8)
    xd = go.vector()
    res = go.linear_problem_solve(xd, method=1, precond=BJ, reduction=1e-10)
    assert res["solver"]["converged"]
    x_ref, res_ref = o.linear_problem_solve(method=1, precond=BJ, reduction=1e-10)
    it, it_ref = res["solver"]["iterations"], res_ref["solver"]["iterations"]
    assert abs(it - it_ref) <= max(2, 0.15 * it_ref)
    x = xd.download()
    r = o.residual(x)
    assert np.linalg.norm(r) <= 1.01e-10 * res["first_defect"]
    assert rel(x, x_ref) <= 1e-7


def test_block_jacobi_transpose_solve(pkg, of):
    say("[gpu] block jacobi transposed BiCGSTAB 5^3 Q2 convection")
    cfg = pkg.make_config(n=5, degree=2, alpha=3.0, b=(1.0, 0.5, 0.25), kappa0=0.1)
    go, o = pkg.GridOperator(cfg), of.Oracle(cfg)
    rhs = np.random.default_rng(22).normal(size=go.N)
    z, r = go.vector(), go.vector(rhs)
    res = go.solve(z, r, method=1, precond=BJ, reduction=1e-12, transpose=True)
    assert res["converged"]
    assert rel(o.jacobian_apply_transpose(z.download()), rhs) <= 2e-12


def test_block_jacobi_mixed_cg(pkg, of):
    say("[gpu] block jacobi mixed CG 10^3 Q3 lognormal kappa")
    n = 10
    cfg = pkg.make_config(n=n, degree=3, alpha=3.0, kappa_per_cell=True, problem=pkg.PROBLEM_CONSTANT, f0=1.0,
                          g0=0.0)
    kappa = np.exp(np.random.default_rng(23).normal(size=n ** 3))
    go, o = pkg.GridOperator(cfg, kappa), of.Oracle(cfg, kappa)
    o.set_threads(8)
    xd = go.vector()
    res = go.solve_mixed(xd, precond=BJ, inner_reduction=0.1, outer_reduction=1e-10, inner_maxit=20000)
    assert res["converged"], res
    x = xd.download()
    x_ref, _ = o.linear_problem_solve(precond=BJ, reduction=1e-13, maxit=20000)
    assert rel(x, x_ref) <= 1e-8
    assert np.linalg.norm(o.residual(x)) <= 1.01e-10 * res["first_defect"]


def test_block_jacobi_follows_kappa(pkg, of):
    """The inverse blocks are rebuilt after set_kappa."""
    say("[gpu] block jacobi after set_kappa 3x4x2 Q2")
    cfg = pkg.make_config(nx=3, ny=4, nz=2, degree=2, alpha=3.0, kappa_per_cell=True)
    rng = np.random.default_rng(24)
    k1, k2 = np.exp(rng.normal(size=24)), np.exp(rng.normal(size=24))
    go, o = pkg.GridOperator(cfg, k1), of.Oracle(cfg, k1)
    z = rng.uniform(-1, 1, go.N)
    y = go.vector()
    go.block_jacobi_apply(go.vector(z), y)
    go.set_kappa(k2)
    o.set_kappa(k2)
    y.zero()
    go.block_jacobi_apply(go.vector(z), y)
    assert rel(y.download(), _solve_blocks(o.block_diagonal(), z, 27)) <= 1e-10
