"""GPU graph reduction (cent_reduce_gpu: degree-1 folding by rounds,
Tarjan-Vishkin biconnected split, packing on the device) against the host
build (cent_reduce, Hopcroft-Tarjan).  Both must give the same biconnected
blocks, the same copy weights and the same in-block edges, compared as sets
since block and copy order may differ; and the reduced closeness / BC
computed from the device build must equal the plain BFS farness (bit-exact)
and the oracle's Brandes (1e-9)."""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


def canon(red):
    """{sorted input vertices of a block: (sorted (vertex, weight) pairs, edge set)}"""
    out = {}
    for b in range(red.nblocks):
        lo, hi = int(red.boff[b]), int(red.boff[b + 1])
        verts = red.porig[lo:hi].tolist()
        key = tuple(sorted(verts))
        assert key not in out, ("two blocks with one vertex set", key[:8])
        wts = tuple(sorted(zip(verts, red.pw[lo:hi].tolist())))
        edges = set()
        for a in range(lo, hi):
            row = red.pcol[red.prp[a]:red.prp[a + 1]]
            assert np.all((row >= lo) & (row < hi)), ("edge leaves its block", b)
            for c in row.tolist():
                edges.add((verts[a - lo], verts[c - lo]))
        out[key] = (wts, frozenset(edges))
    return out


def check_same(ctx, g, what, bc=False):
    print(f"[gpu] device reduction {what}: n={g.n} m={g.m}", flush=True)
    host = bc_amd.Reduced(g)
    dev = bc_amd.Reduced(g, ctx=ctx)
    for f in ("n", "np", "nnzp", "nblocks", "nfold", "n_ap", "ncomp"):
        assert getattr(dev, f) == getattr(host, f), (what, f, getattr(dev, f), getattr(host, f))
    # rows sorted and strictly increasing (no duplicate edges)
    for a in range(dev.np):
        row = dev.pcol[dev.prp[a]:dev.prp[a + 1]]
        assert np.all(np.diff(row) > 0), (what, "row not sorted", a)
    assert canon(dev) == canon(host), (what, "blocks differ")
    want = O.weighted_farness(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64))
    for sched in (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_PER_ITEM, bc_amd.CENT_SCHED_SHARED):
        got = ctx.cc_reduced(dev, sched)
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, (what, sched, bad[:10], got[bad[:5]], want[bad[:5]])
    if bc:
        wbc, wfar, _, _ = O.brandes(g, None, threads=min(16, O.max_threads()))
        got_bc, got_far = ctx.bc_reduced(dev)
        np.testing.assert_allclose(got_bc, wbc, rtol=1e-9, atol=1e-9 * max(1.0, float(np.abs(wbc).max())),
                                   err_msg=what)
        assert np.array_equal(got_far, wfar), what
    return host, dev


def path(n):
    return bc_amd.CSR.from_edges(n, np.arange(n - 1), np.arange(1, n))


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt",
                                  "testdata.txt", "testdata1.txt", "testdata2.txt", "ia-fb-messages.mtx"])
def test_device_reduction_fixtures(ctx, name):
    check_same(ctx, bc_amd.buildGraph(O.dataset_path(name)), name, bc=True)


def test_device_reduction_trees_and_cycles(ctx):
    rng = np.random.default_rng(5)
    # a path folds one leaf from each end per round: 150 rounds
    check_same(ctx, path(300), "path300", bc=True)
    # star, K2 pieces and isolated vertices: every component folds to one vertex
    u = np.array([0, 0, 0, 0, 5, 7], np.int64)
    v = np.array([1, 2, 3, 4, 6, 8], np.int64)
    check_same(ctx, bc_amd.CSR.from_edges(12, u, v), "star+K2+isolated", bc=True)
    # random tree
    n = 1000
    par = np.array([rng.integers(0, i) for i in range(1, n)])
    check_same(ctx, bc_amd.CSR.from_edges(n, np.arange(1, n), par), "random-tree", bc=True)
    # cycle, and two triangles sharing a vertex (one articulation point)
    check_same(ctx, bc_amd.CSR.from_edges(7, np.arange(7), (np.arange(7) + 1) % 7), "cycle7", bc=True)
    check_same(ctx, bc_amd.CSR.from_edges(5, np.array([0, 1, 2, 2, 3, 4]), np.array([1, 2, 0, 3, 4, 2])),
               "bowtie", bc=True)
    # cycles with pendant trees, chains of bridges and several components
    edges = []
    base = 0
    for k in range(6):
        c = 3 + 2 * k
        edges += [(base + i, base + (i + 1) % c) for i in range(c)]
        t = base + c
        for i in range(5 + k):  # pendant tree hanging off vertex base
            edges.append((base + int(rng.integers(0, c)) if i == 0 else t + int(rng.integers(0, i)), t + i))
        base = t + 5 + k + 3  # 3 isolated vertices between components
    e = np.array(edges)
    check_same(ctx, bc_amd.CSR.from_edges(base, e[:, 0], e[:, 1]), "cycles+pendants", bc=True)


def test_device_reduction_block_trees(ctx):
    check_same(ctx, bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=3), "blocktree-300", bc=True)
    check_same(ctx, bc_amd.gen_blocktree(6, 1.5, 2500, 6000, 6.0, 0.3, 0.5, seed=8), "blocktree-big")
    check_same(ctx, bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3), "blocktree-2000")


def test_device_reduction_power_law_and_grid(ctx):
    check_same(ctx, bc_amd.gen_rmat(13, 8 << 13, seed=12, lcc=False), "rmat13-all-components", bc=True)
    check_same(ctx, bc_amd.gen_grid(40, 40, 0.1, 4, seed=4), "grid40", bc=True)
    check_same(ctx, bc_amd.buildGraph(O.dataset_path("musae_git.txt")), "musae_git")


def test_device_reduction_large(ctx):
    # R-MAT with hubs (one giant block) and a deep, wide block tree: closeness
    # from the device build equals the host build's (both are pinned to plain
    # farness on the smaller graphs above)
    for g, what in ((bc_amd.gen_rmat(16, 16 << 16, seed=7, lcc=False), "rmat16"),
                    (bc_amd.gen_blocktree(20000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=9), "blocktree-20000")):
        print(f"[gpu] device reduction {what}: n={g.n} m={g.m}", flush=True)
        host = bc_amd.Reduced(g)
        dev = bc_amd.Reduced(g, ctx=ctx)
        for f in ("np", "nnzp", "nblocks", "nfold", "n_ap", "ncomp"):
            assert getattr(dev, f) == getattr(host, f), (what, f)
        assert np.array_equal(ctx.cc_reduced(dev), ctx.cc_reduced(host)), what


def test_device_reduction_repeat_and_empty(ctx):
    g = bc_amd.gen_blocktree(200, 1.5, 8, 3000, 6.0, 0.3, 0.5, seed=11)
    a = bc_amd.Reduced(g, ctx=ctx)
    b = bc_amd.Reduced(g, ctx=ctx)
    for f in ("prp", "pcol", "pw", "porig", "boff"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f  # deterministic layout
    one = bc_amd.CSR.from_edges(1, np.zeros(0, np.int64), np.zeros(0, np.int64))
    r = bc_amd.Reduced(one, ctx=ctx)
    assert r.np == 0 and r.nblocks == 0 and r.n == 1
    assert ctx.cc_reduced(r).tolist() == [0]
