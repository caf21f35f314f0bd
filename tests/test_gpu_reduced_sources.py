"""Sampled-source betweenness through the D1/AP reduction on the GPU
(cent_bc_reduced_sources) against the oracle's plain Brandes over the same
sources (Seq_Brandes/Brandes.c:224-372 restated in FP64) and against cent_bc:
BC within 1e-9 relative, farness of every source bit-exact, under both
schedules (persistent items with source lists, level-synchronous ranges)."""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu

SCHEDS = (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_PER_ITEM)


def check(ctx, g, src, what, plain=True):
    print(f"[gpu] reduced sources {what}: n={g.n} k={len(src)}", flush=True)
    red = bc_amd.Reduced(g)
    obc, ofar, _, _ = O.brandes(g, src, threads=min(16, O.max_threads()))
    scale = max(1.0, float(np.abs(obc).max()))
    for sched in SCHEDS:
        bc, far = ctx.bc_reduced_sources(red, src, sched)
        assert np.array_equal(far, ofar), (what, sched, np.nonzero(far != ofar)[0][:8])
        err = float(np.max(np.abs(bc - obc))) / scale
        assert err <= 1e-9, (what, sched, err)
    if plain:
        pbc, pfar, _ = ctx.bc(ctx.upload(g), src)
        assert np.array_equal(pfar, far)
        assert float(np.max(np.abs(pbc - bc))) / scale <= 1e-9
    # the GPU-built reduction gives the same scores
    bc2, far2 = ctx.bc_reduced_sources(bc_amd.Reduced(g, ctx=ctx), src)
    assert np.array_equal(far2, ofar) and float(np.max(np.abs(bc2 - obc))) / scale <= 1e-9


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "testdata2.txt", "musae_git.txt"])
def test_fixtures(ctx, name):
    g = bc_amd.buildGraph(O.dataset_path(name))
    rng = np.random.default_rng(len(name))
    check(ctx, g, rng.integers(g.startNodeID, g.n, 40).astype(np.int32), name)  # repeats included


def test_block_tree_workload_sizes(ctx):
    g = bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3)
    for k, seed in ((64, 1), (700, 2)):
        check(ctx, g, bc_amd.sample_sources(g.n, k, seed), f"blocktree2000 k={k}")


def test_giant_block_and_deep_small_block(ctx):
    # R-MAT: one block far beyond the LDS window (level-synchronous ranges)
    g = bc_amd.gen_rmat(13, 16 << 13, seed=9)
    check(ctx, g, bc_amd.sample_sources(g.n, 150, 3), "rmat13")
    # a 700-cycle (a block deeper than the item kernel's level limit) with pendant paths and a bowtie
    n = 700
    u = list(range(n)) + [0, 701, 702, 350, 704, 705, 706, 704]
    v = [(i + 1) % n for i in range(n)] + [700, 700, 701, 703, 703, 704, 705, 706]
    g = bc_amd.CSR.from_edges(708, u, v)
    check(ctx, g, np.array([0, 5, 350, 350, 702, 706, 707], np.int32), "cycle700")


def test_empty_and_all_sources(ctx):
    g = bc_amd.gen_blocktree(80, 1.5, 8, 300, 6.0, 0.3, 0.5, seed=5)
    red = bc_amd.Reduced(g)
    bc, far = ctx.bc_reduced_sources(red, np.zeros(0, np.int32))
    assert np.all(bc == 0) and len(far) == 0
    allb, allf = ctx.bc_reduced_sources(red, np.arange(g.n, dtype=np.int32))
    rb, rf = ctx.bc_reduced(red)
    assert np.array_equal(allf, rf)
    assert float(np.max(np.abs(allb - rb))) / max(1.0, float(np.abs(rb).max())) <= 1e-9
    with pytest.raises(bc_amd.CentError):
        ctx.bc_reduced_sources(red, np.array([g.n], np.int32))
