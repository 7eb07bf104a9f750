"""The allocation-free L2 error entry point (pdelab_gpu_l2_error_squared_noalloc) against
the allocating one, through the C-ABI: bit-identical results on two calls in a row, the
second on a smaller grid (the shared scratch array keeps the first call's partials, which
the second must not pick up), and a clear error above the scratch maximum."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(pkg, n, degree, seed, order):
    cfg = pkg.make_config(nx=n[0], ny=n[1], nz=n[2], degree=degree, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    x = go.vector(np.random.default_rng(seed).uniform(-1.0, 1.0, go.N))
    ref = go.l2_error_squared(x, order)
    got = go.l2_error_squared_noalloc(x, order)
    return go, ref, got


def test_noalloc_l2_error_matches_allocating_entry_point(pkg):
    say("[gpu] no-alloc l2 error: 40x36x32 Q3 then 6x5x4 Q2")
    go1, ref1, got1 = _case(pkg, (40, 36, 32), 3, 1, 10)
    assert got1 == ref1 and ref1 > 0.0
    go2, ref2, got2 = _case(pkg, (6, 5, 4), 2, 2, 7)
    assert got2 == ref2 and ref2 > 0.0
    # the first grid again, after the smaller call
    x1 = go1.vector(np.random.default_rng(1).uniform(-1.0, 1.0, go1.N))
    assert go1.l2_error_squared_noalloc(x1, 10) == ref1
    go1.close()
    go2.close()


def test_noalloc_l2_error_rejects_grids_above_the_maximum(pkg):
    say("[gpu] no-alloc l2 error: grid above the scratch maximum")
    n = 102  # 102^3 = 1,061,208 cells > 2^20
    assert n ** 3 > pkg.NOALLOC_MAX_CELLS
    cfg = pkg.make_config(n=n, degree=1, alpha=3.0, problem=pkg.PROBLEM_EXP)
    go = pkg.GridOperator(cfg)
    x = go.vector()
    with pytest.raises(pkg.PdelabError) as e:
        go.l2_error_squared_noalloc(x, 4)
    assert e.value.status == 1 and "scratch maximum" in str(e.value)
    assert go.l2_error_squared(x, 4) > 0.0  # the allocating entry point takes it
    go.close()
