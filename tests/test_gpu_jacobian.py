"""GPU parity of the assembled Jacobian (7-point block pattern) and its block-matrix apply.

* pdelab_gpu_jacobian against the oracle's jacobian_volume/skeleton/boundary assembly
  (R:dune/pdelab/gridoperator/default/assembler.hh:84-279 restated): same column pattern,
  blocks within 1e-12 relative, for every operator variant of the apply parity cases; the
  transposed assembly against the block transpose;
* pdelab_gpu_block_matrix_apply against the matrix-free operator apply (the identity the
  reference's matrix_free_linear.cc relies on): within 1e-12 relative.
"""
import numpy as np
import pytest

from test_gpu_parity import CASES, make, rel, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,kw,seed", CASES, ids=[c[0] for c in CASES])
def test_jacobian_blocks_and_apply(pkg, of, name, kw, seed):
    say(f"[gpu] jacobian {name}")
    cfg, kappa, go, o = make(pkg, of, kw, seed)
    nloc = (cfg.degree + 1) ** 3
    ref_blocks, ref_cols = o.assemble()
    bd, cd = go.jacobian()
    blocks = bd.download().reshape(ref_blocks.shape)
    cols = cd.download().view(np.int64).reshape(ref_cols.shape)
    assert (cols == ref_cols).all()
    assert rel(blocks, ref_blocks) <= 1e-12
    # matrix-based apply == matrix-free apply
    z = np.random.default_rng(31).uniform(-1, 1, go.N)
    zd, yd, ym = go.vector(z), go.vector(np.full(go.N, 7.0)), go.vector()
    go.block_matrix_apply(bd, cd, zd, yd)
    go.operator_apply(zd, ym)
    assert rel(yd.download(), ym.download()) <= 1e-12
    # J^T: block (c, s) is the transpose of block (neighbour, opposite face) of J
    bt, ct = go.jacobian(transpose=True)
    blocks_t = bt.download().reshape(ref_blocks.shape)
    assert (ct.download().view(np.int64).reshape(ref_cols.shape) == ref_cols).all()
    want = np.zeros_like(ref_blocks)
    want[:, 0] = np.transpose(ref_blocks[:, 0], (0, 2, 1))
    for c in range(go.ncells):
        for f in range(6):
            cn = ref_cols[c, 1 + f]
            if cn >= 0:
                want[c, 1 + f] = ref_blocks[cn, 1 + (f ^ 1)].T
    assert rel(blocks_t, want) <= 1e-12
    del nloc
