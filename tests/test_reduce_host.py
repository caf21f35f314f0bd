"""Host reduction (degree-1 folding + biconnected-block split + block-cut tree
constants) checked on the CPU: the packed weighted distance sums D are
computed by the oracle, cent_reduced_farness turns them into farness, and
the result must equal plain BFS farness exactly (int64) on every graph —
connected or not, trees, cycles, block trees, power-law graphs."""
import itertools

import numpy as np
import pytest

import _oracle as O
import bc_amd


def reduced_farness_cpu(g):
    r = bc_amd.Reduced(g)
    D = O.weighted_farness(r.np, r.prp, r.pcol, r.pw)
    return r, r.farness_from(D)


def check(g):
    r, far = reduced_farness_cpu(g)
    want = O.farness(g)
    assert np.array_equal(far, want), np.nonzero(far != want)[0][:10]
    # every block's weights add up to its component size
    for b in range(r.nblocks):
        lo, hi = r.boff[b], r.boff[b + 1]
        assert hi > lo
    return r


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt",
                                  "testdata.txt", "testdata1.txt", "testdata2.txt",
                                  "ia-fb-messages.mtx"])
def test_fixtures(name):
    check(bc_amd.buildGraph(O.dataset_path(name)))


def test_reference_known_answers_through_reduction():
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    _, far = reduced_farness_cpu(g)
    assert far[7] == 86
    g = bc_amd.buildGraph(O.dataset_path("musae_git.txt"))
    r, far = reduced_farness_cpu(g)
    assert far[7] == 117726
    assert r.nfold > 0 and r.n_ap > 0


def path(n):
    return bc_amd.CSR.from_edges(n, np.arange(n - 1), np.arange(1, n))


def test_shapes():
    check(path(2))
    check(path(17))                                       # tree: folds to one vertex
    star = bc_amd.CSR.from_edges(9, np.zeros(8), np.arange(1, 9))
    check(star)
    cyc = bc_amd.CSR.from_edges(10, np.arange(10), (np.arange(10) + 1) % 10)
    r = check(cyc)
    assert r.nfold == 0 and r.nblocks == 1
    # two triangles joined by a path (bridges), with pendant chains
    u = [0, 1, 2, 2, 3, 4, 5, 6, 6, 4, 7, 8]
    v = [1, 2, 0, 3, 4, 5, 6, 7, 5, 9, 9, 10]
    check(bc_amd.CSR.from_edges(11, u, v))
    # bowtie: AP shared by two cycles + isolated vertex + isolated edge
    u = [0, 1, 2, 2, 3, 4, 6]
    v = [1, 2, 0, 3, 4, 2, 7]
    check(bc_amd.CSR.from_edges(9, u, v))
    u, v = zip(*itertools.combinations(range(6), 2))
    check(bc_amd.CSR.from_edges(6, u, v))


@pytest.mark.parametrize("seed", range(12))
def test_random_sparse_graphs(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 80))
    m = int(rng.integers(n // 2, 2 * n))
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    check(bc_amd.CSR.from_edges(n, u, v))


def test_block_tree_and_rmat():
    t = bc_amd.gen_blocktree(60, 1.5, 8, 400, 6.0, 0.3, 0.5, seed=3)
    r = check(t)
    assert r.n_ap > 10 and r.nfold > 0
    g = bc_amd.gen_rmat(11, 8 << 11, seed=7, lcc=False)
    r = check(g)
    assert r.ncomp > 1
    grid = bc_amd.gen_grid(20, 20, 0.1, 5, seed=4)
    check(grid)


def check_bc(g):
    """Reduced betweenness (combinatorial piece term + weighted in-block Brandes
    from the oracle) must equal plain all-source Brandes."""
    r = bc_amd.Reduced(g)
    bcp, D = O.weighted_brandes(r.np, r.prp, r.pcol, r.pw)
    bc = r.bc_from(bcp)
    want, far, _, _ = O.brandes(g, None)
    np.testing.assert_allclose(bc, want, rtol=1e-10, atol=1e-9 * max(1.0, np.abs(want).max()))
    assert np.array_equal(r.farness_from(D), far)


def test_weighted_oracle_reduces_to_plain():
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    bc, far = O.weighted_brandes(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64))
    want, wfar, _, _ = O.brandes(g, None)
    np.testing.assert_allclose(bc, want, rtol=1e-12)
    assert np.array_equal(far, wfar)
    # listed sources with their own weights: unit weights give plain Brandes
    # over the list, and a source listed with weight 2 counts twice
    src = np.array([3, 9, 9, 20], np.int32)
    bc2, far2 = O.weighted_brandes_src(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64), src, np.ones(4, np.int64))
    want2, wfar2, _, _ = O.brandes(g, src)
    np.testing.assert_allclose(bc2, want2, rtol=1e-12)
    assert np.array_equal(far2, wfar2)
    bc3, _ = O.weighted_brandes_src(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64), src[[0, 1, 3]],
                                    np.array([1, 2, 1], np.int64))
    np.testing.assert_allclose(bc3, want2, rtol=1e-12)


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt",
                                  "testdata.txt", "testdata1.txt", "testdata2.txt",
                                  "ia-fb-messages.mtx"])
def test_reduced_bc_fixtures(name):
    check_bc(bc_amd.buildGraph(O.dataset_path(name)))


def test_reduced_bc_shapes_and_random():
    check_bc(path(2))
    check_bc(path(15))
    check_bc(bc_amd.CSR.from_edges(9, np.zeros(8), np.arange(1, 9)))
    u = [0, 1, 2, 2, 3, 4, 5, 6, 6, 4, 7, 8]
    v = [1, 2, 0, 3, 4, 5, 6, 7, 5, 9, 9, 10]
    check_bc(bc_amd.CSR.from_edges(11, u, v))
    u = [0, 1, 2, 2, 3, 4, 6]
    v = [1, 2, 0, 3, 4, 2, 7]
    check_bc(bc_amd.CSR.from_edges(9, u, v))
    for seed in range(10):
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(5, 70))
        m = int(rng.integers(n // 2, 2 * n))
        check_bc(bc_amd.CSR.from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m)))
    check_bc(bc_amd.gen_blocktree(40, 1.5, 8, 200, 6.0, 0.3, 0.5, seed=3))
    check_bc(bc_amd.gen_rmat(9, 8 << 9, seed=7, lcc=False))


# ---- sampled sources through the reduction (the host side of cent_bc_reduced_sources)
def _sampled_check(g, src):
    r = bc_amd.Reduced(g)
    WS = r.source_weights(src)
    T = np.nonzero(WS > 0)[0].astype(np.int32)
    bcp, _ = O.weighted_brandes_src(r.np, r.prp, r.pcol, r.pw, T, WS[T])
    got = r.bc_from_sources(src, bcp)
    want, wfar, _, _ = O.brandes(g, src)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(want).max()))
    D = O.weighted_farness(r.np, r.prp, r.pcol, r.pw)
    assert np.array_equal(r.farness_of_sources(src, D), wfar)
    return r, WS, T


def test_source_weights_of_all_vertices_are_copy_weights():
    for g in (bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=3), bc_amd.gen_rmat(11, 16 << 11, seed=4),
              bc_amd.buildGraph(O.dataset_path("testdata2.txt"))):
        r = bc_amd.Reduced(g)
        assert np.array_equal(r.source_weights(np.arange(g.n)), r.pw)


@pytest.mark.parametrize("name", ["karate.txt", "dolphins.txt", "testdata2.txt", "football.txt"])
def test_sampled_reduced_bc_fixtures(name):
    g = bc_amd.buildGraph(O.dataset_path(name))
    rng = np.random.default_rng(len(name))
    for k in (1, 5, 17):
        src = rng.integers(g.startNodeID, g.n, k).astype(np.int32)  # repeats allowed
        _sampled_check(g, src)
    _sampled_check(g, g.sources())  # all sources: the all-source formula


def test_sampled_reduced_bc_block_trees_and_forests():
    g = bc_amd.gen_blocktree(400, 1.5, 8, 3000, 6.0, 0.3, 0.5, seed=11)
    for k, seed in ((64, 1), (300, 2), (7, 3)):
        src = bc_amd.sample_sources(g.n, k, seed)
        r, WS, T = _sampled_check(g, src)
        assert len(T) < r.np // 2  # far fewer traversals than all-source
    # a forest of pendant trees and paths (everything folds), a cycle, isolated vertices
    u = [0, 1, 1, 3, 5, 6, 7, 8, 9, 11]
    v = [1, 2, 3, 4, 6, 7, 8, 5, 10, 12]
    g = bc_amd.CSR.from_edges(15, u, v)
    for src in ([0], [2, 2, 4], [5, 9, 13], list(range(15))):
        _sampled_check(g, np.array(src, np.int32))
