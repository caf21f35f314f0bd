"""GPU parity of the reduced (degree-1 folding + articulation split) closeness
path against plain BFS farness — the invariance that pins it: both schedules
(persistent work queue, per-item launches) must give the plain farness of
every vertex, bit-exact in int64."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SCHEDS = (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_PER_ITEM)
# closeness also through the neighbour-sharing schedule (the SI/R rule)
CC_SCHEDS = SCHEDS + (bc_amd.CENT_SCHED_SHARED,)


def plain_farness(g):
    return O.weighted_farness(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64))


def check(ctx, g, what):
    print(f"[gpu] reduced {what}: n={g.n} m={g.m}", flush=True)
    red = bc_amd.Reduced(g)
    want = plain_farness(g)
    for sched in CC_SCHEDS:
        got = ctx.cc_reduced(red, sched)
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, (what, sched, bad[:10], got[bad[:5]], want[bad[:5]])
    return red


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt",
                                  "testdata.txt", "testdata1.txt", "testdata2.txt",
                                  "ia-fb-messages.mtx"])
def test_reduced_fixtures(ctx, name):
    check(ctx, bc_amd.buildGraph(O.dataset_path(name)), name)


def test_reference_shim_and_known_answers(ctx):
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    CCs = np.zeros(g.n, np.int64)
    bc_amd.compute_D1_AP_CC_shareBased_DegreeOrder(g, CCs)
    assert CCs[7] == 86
    ref = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))["karate.txt"]
    assert CCs[1:35].tolist() == ref["d1ap"]
    g = bc_amd.buildGraph(O.dataset_path("musae_git.txt"))
    red = check(ctx, g, "musae_git")
    far = ctx.cc_reduced(red)
    assert far[7] == 117726


def test_reduced_block_trees(ctx):
    # many small blocks (persistent queue) + pendant chains
    check(ctx, bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=3), "blocktree-300")
    # blocks larger than the LDS item range (level-synchronous path)
    check(ctx, bc_amd.gen_blocktree(6, 1.5, 2500, 6000, 6.0, 0.3, 0.5, seed=8), "blocktree-big")


def test_reduced_power_law_and_grid(ctx):
    # giant block with hub rows + many components, trees and pendants
    check(ctx, bc_amd.gen_rmat(13, 8 << 13, seed=12, lcc=False), "rmat13-all-components")
    check(ctx, bc_amd.gen_grid(40, 40, 0.1, 4, seed=4), "grid40")


def check_bc(ctx, g, what):
    print(f"[gpu] reduced bc {what}: n={g.n} m={g.m}", flush=True)
    red = bc_amd.Reduced(g)
    want, wfar, _, _ = O.brandes(g, None, threads=min(16, O.max_threads()))
    scale = max(1.0, float(np.abs(want).max()))
    out = {}
    for sched in SCHEDS:
        bc, far = ctx.bc_reduced(red, sched)
        np.testing.assert_allclose(bc, want, rtol=1e-9, atol=1e-9 * scale, err_msg=f"{what} sched {sched}")
        assert np.array_equal(far, wfar), (what, sched)
        out[sched] = bc
    return red, out


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt",
                                  "testdata.txt", "testdata1.txt", "testdata2.txt",
                                  "ia-fb-messages.mtx", "musae_git.txt"])
def test_reduced_bc_fixtures(ctx, name):
    check_bc(ctx, bc_amd.buildGraph(O.dataset_path(name)), name)


def test_reduced_bc_shapes(ctx):
    # long cycle: one block deeper than the persistent kernel's level limit
    n = 600
    check_bc(ctx, bc_amd.CSR.from_edges(n, np.arange(n), (np.arange(n) + 1) % n), "cycle600")
    check_bc(ctx, bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=3), "blocktree-300")
    check_bc(ctx, bc_amd.gen_blocktree(6, 1.5, 2500, 6000, 6.0, 0.3, 0.5, seed=8), "blocktree-big")
    check_bc(ctx, bc_amd.gen_rmat(12, 8 << 12, seed=12, lcc=False), "rmat12-all-components")
    check_bc(ctx, bc_amd.gen_grid(30, 30, 0.1, 4, seed=4), "grid30")


def test_reduced_bc_deterministic(ctx):
    g = bc_amd.gen_blocktree(200, 1.5, 8, 3000, 6.0, 0.3, 0.5, seed=11)
    red = bc_amd.Reduced(g)
    a = ctx.bc_reduced(red)
    b = ctx.bc_reduced(red)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_shared_schedule_shapes(ctx):
    # centers with more than 32 unanswered neighbours (a star inside a block),
    # with exactly 3, and blocks where every source stays alone (a cycle);
    # BC through the shared schedule is refused
    n = 80
    hub = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, n - 1)]  # a fan: one block
    e = np.array(hub)
    check(ctx, bc_amd.CSR.from_edges(n, e[:, 0], e[:, 1]), "fan80")
    check(ctx, bc_amd.CSR.from_edges(9, np.arange(9), (np.arange(9) + 1) % 9), "cycle9")
    k4 = np.array([(a, b) for a in range(4) for b in range(a + 1, 4)])
    check(ctx, bc_amd.CSR.from_edges(4, k4[:, 0], k4[:, 1]), "K4")
    red = bc_amd.Reduced(bc_amd.CSR.from_edges(4, k4[:, 0], k4[:, 1]))
    with pytest.raises(bc_amd.CentError):
        ctx.bc_reduced(red, bc_amd.CENT_SCHED_SHARED)
