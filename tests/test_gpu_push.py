"""The dense betweenness step's push/pull switch at level 2.

A batch's level 2 is either pulled (every row scanned against the level-1
masks) or pushed from the level-1 frontier (k_push2_*: OR of the level-2
masks, the finishing pass, then the path counts added along the same
edges).  Level-1 path counts are 1 per lane, so the level-2 counts are
exact integers in either order: the two forms must agree bit for bit, and
both with the oracle (Seq_Brandes/Brandes.c:224-372 in FP64: farness and
distance sums exact, BC within 1e-9 relative).
"""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


def _graphs():
    yield "rmat14", bc_amd.gen_rmat(14, 8 << 14, seed=1)
    yield "er15", bc_amd.gen_rmat(15, 16 << 15, 0.25, 0.25, 0.25, seed=3)
    yield "karate", bc_amd.buildGraph(O.dataset_path("karate.txt"))


@pytest.mark.parametrize("mode", [bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT])
def test_level2_push_equals_pull_and_oracle(ctx, mode):
    for name, g in _graphs():
        dg = ctx.upload(g)
        src = g.sources() if g.n < 100 else bc_amd.sample_sources(g.n, 200, 5)
        print(f"[gpu] {name} n={g.n}: level 2 pulled vs pushed, mode {mode}", flush=True)
        pull = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier="dense", push_pct=-1)
        push = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier="dense", push_pct=1000)
        for a, b in zip(pull, push):
            assert np.array_equal(a, b), name
        obc, ofar, ods, _ = O.brandes(g, src, mode=mode)
        assert np.array_equal(push[1], ofar) and np.array_equal(push[2], ods), name
        scale = max(1.0, float(np.abs(obc).max()))
        np.testing.assert_allclose(push[0], obc, rtol=1e-9, atol=1e-9 * scale, err_msg=name)


def test_level2_push_through_the_reduction(ctx):
    """BC through the D1/AP reduction runs its large blocks on the level
    kernels with vertex weights (per-item launches: every block): pushed and
    pulled level 2 give the same bits, and the oracle's results."""
    g = bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=7)
    red = bc_amd.Reduced(g)
    obc, ofar, _, _ = O.brandes(g, None)
    pull = ctx.bc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM, push_pct=-1)
    push = ctx.bc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM, push_pct=1000)
    assert np.array_equal(pull[0], push[0]) and np.array_equal(pull[1], push[1])
    assert np.array_equal(push[1], ofar)
    scale = max(1.0, float(np.abs(obc).max()))
    np.testing.assert_allclose(push[0], obc, rtol=1e-9, atol=1e-9 * scale)
