"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake (a dropped term, an off-by-one
distance, the wrong accumulation formula) fails at least one of them:
reference known answers and golden vectors, outputs of the reference's own
code built from its sources, networkx, closed forms and brute force.
"""
import itertools
import json
import os
from collections import deque

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN


def load(name):
    return bc_amd.buildGraph(O.dataset_path(name))


# ---------------------------------------------------------------- brute force
def bfs_all(n, adj, s):
    dist = [-1] * n
    sig = [0] * n
    dist[s], sig[s] = 0, 1
    q = deque([s])
    while q:
        x = q.popleft()
        for y in adj[x]:
            if dist[y] < 0:
                dist[y] = dist[x] + 1
                q.append(y)
            if dist[y] == dist[x] + 1:
                sig[y] += sig[x]
    return dist, sig


def brute_bc(g, compat=False):
    """Pair-dependency definition: BC(v) = sum_{s != v != t} sigma_st(v)/sigma_st,
    with sigma_st(v) = sigma_sv * sigma_vt when d(s,v) + d(v,t) = d(s,t).  In the
    compat form each source's dependency delta_s(v) is weighted by the number
    of shortest-path predecessors of v from s."""
    n = g.n
    adj = [g.col[g.row_ptr[v]:g.row_ptr[v + 1]].tolist() for v in range(n)]
    D, S = zip(*[bfs_all(n, adj, s) for s in range(n)])
    bc = np.zeros(n)
    for s in range(n):
        for v in range(n):
            if v == s or D[s][v] < 0:
                continue
            delta = 0.0
            for t in range(n):
                if t in (s, v) or D[s][t] < 0 or D[v][t] < 0:
                    continue
                if D[s][v] + D[v][t] == D[s][t]:
                    delta += S[s][v] * S[v][t] / S[s][t]
            if compat:
                npred = sum(1 for u in adj[v] if D[s][u] >= 0 and D[s][u] + 1 == D[s][v])
                delta *= npred
            bc[v] += delta
    far = np.array([sum(d for d in D[s] if d > 0) for s in range(n)], np.int64)
    return bc, far


def random_graph(n, p, seed):
    rng = np.random.default_rng(seed)
    u, v = np.nonzero(np.triu(rng.random((n, n)) < p, 1))
    return bc_amd.CSR.from_edges(n, u, v)


@pytest.mark.parametrize("seed", range(6))
def test_brandes_matches_brute_force(seed):
    g = random_graph(9 + seed, 0.3, seed)
    for mode in (0, 1):
        bc, far, _, _ = O.brandes(g, mode=mode)
        bbc, bfar = brute_bc(g, compat=(mode == 1))
        np.testing.assert_allclose(bc, bbc, rtol=1e-12, atol=1e-12)
        assert np.array_equal(far, bfar)
        assert np.array_equal(O.farness(g), bfar)


# ---------------------------------------------------------------- closed forms
def path(n):
    return bc_amd.CSR.from_edges(n, np.arange(n - 1), np.arange(1, n))


def test_closed_forms():
    n = 12
    bc, far, _, _ = O.brandes(path(n))
    i = np.arange(n)
    np.testing.assert_allclose(bc, 2 * i * (n - 1 - i), rtol=0, atol=1e-9)
    assert np.array_equal(far, i * (i + 1) // 2 + (n - 1 - i) * (n - i) // 2)
    star = bc_amd.CSR.from_edges(n, np.zeros(n - 1), np.arange(1, n))
    bc, far, _, _ = O.brandes(star)
    assert bc[0] == (n - 1) * (n - 2) and np.all(bc[1:] == 0)
    assert far[0] == n - 1 and np.all(far[1:] == 1 + 2 * (n - 2))
    cyc = bc_amd.CSR.from_edges(n, np.arange(n), (np.arange(n) + 1) % n)
    bc, far, _, _ = O.brandes(cyc)
    assert np.allclose(bc, bc[0]) and np.all(far == far[0]) and far[0] == (n // 2) ** 2
    # uniform BC on a vertex-transitive graph: sum identity fixes the value
    assert np.isclose(bc[0] * n, far.sum() - n * (n - 1))
    u, v = zip(*itertools.combinations(range(7), 2))
    kg = bc_amd.CSR.from_edges(7, u, v)
    bc, far, _, _ = O.brandes(kg)
    assert np.all(bc == 0) and np.all(far == 6)


def test_full_grid_farness_is_manhattan():
    r, c = 7, 9
    ids = np.arange(r * c).reshape(r, c)
    u = np.concatenate([ids[:, :-1].ravel(), ids[:-1, :].ravel()])
    v = np.concatenate([ids[:, 1:].ravel(), ids[1:, :].ravel()])
    g = bc_amd.CSR.from_edges(r * c, u, v)
    far = O.farness(g)
    y, x = np.divmod(np.arange(r * c), c)
    expect = np.abs(y[:, None] - y[None, :]).sum(1) + np.abs(x[:, None] - x[None, :]).sum(1)
    assert np.array_equal(far, expect)


# ---------------------------------------------------------------- reference answers
def test_reference_closeness_known_answers():
    # Lib/AP_Process/AP_process_karate.txt:19 and AP_process_musae_git.txt:19
    g = load("karate.txt")
    assert O.farness(g, [7])[0] == 86
    g = load("musae_git.txt")
    assert O.farness(g, [7])[0] == 117726


def test_reference_karate_bc_score_file():
    # karate_BC_Score.txt: header "1 34", one FP32 score per id, per-predecessor formula
    lines = open(os.path.join(GOLDEN, "karate_BC_Score.txt")).read().split()
    start, end = int(lines[0]), int(lines[1])
    ref = np.array([float(x) for x in lines[2:]])
    g = load("karate.txt")
    src = g.sources()
    assert (start, end) == (g.startNodeID, g.endNodeID)
    bc1, _, _, _ = O.brandes(g, src, mode=1)
    np.testing.assert_allclose(bc1[start:end + 1], ref, rtol=2e-6, atol=2e-6)
    f32 = O.brandes_f32(g, src)
    np.testing.assert_allclose(f32[start:end + 1], ref, rtol=2e-6, atol=2e-6)
    # the textbook formula does NOT reproduce the file (vertex 1: 462.14 vs 792.49)
    bc0, _, _, _ = O.brandes(g, src, mode=0)
    assert abs(bc0[1] - 462.142857) < 1e-5 and abs(bc0[1] - ref[0]) > 100


def test_oracle_matches_reference_build_outputs():
    """ref_golden.json holds the outputs of the reference's own computeBC,
    computeCC and compute_D1_AP_CC_shareBased_DegreeOrder (built from its sources
    by oracle/Makefile, written by tests/golden/make_golden.py)."""
    ref = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))
    for name, e in ref.items():
        if e.get("raw"):
            continue  # multigraph fixtures: tests/test_raw_input.py
        g = load(name)
        src = g.sources()
        s = e["start"]
        far = O.farness(g, src)
        assert np.array_equal(far, np.array(e["cc"])), name
        assert np.array_equal(far, np.array(e["d1ap"])), name
        bc1, _, _, _ = O.brandes(g, src, mode=1)
        # FP64 vs the reference FP32 sums over up to ~1.3k sources: FP32 rounding bound
        np.testing.assert_allclose(bc1[s:s + len(src)], e["bc_f32"], rtol=1e-4, atol=5e-5, err_msg=name)


def test_oracle_matches_networkx_golden():
    gold = json.load(open(os.path.join(GOLDEN, "nx_golden.json")))
    for name, e in gold.items():
        g = load(name)
        src = g.sources()
        s = e["start"]
        bc0, far, _, _ = O.brandes(g, src, mode=0)
        np.testing.assert_allclose(bc0[s:s + len(src)], e["bc"], rtol=1e-9, atol=1e-9, err_msg=name)
        assert np.array_equal(far, np.array(e["farness"])), name


def test_sum_identity_on_connected_graphs():
    for name in ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt"]:
        g = load(name)
        src = g.sources()
        bc, far, _, dag = O.brandes(g, src)
        nn = len(src)
        assert np.isclose(bc.sum(), far.sum() - nn * (nn - 1), rtol=1e-12)
        assert dag >= nn - 1


def test_planted_mistake_is_caught():
    """A wrong oracle (farness counted from 1 instead of 0) must fail the identity."""
    g = load("karate.txt")
    src = g.sources()
    bc, far, _, _ = O.brandes(g, src)
    wrong = far + (len(src) - 1)
    assert not np.isclose(bc.sum(), wrong.sum() - len(src) * (len(src) - 1), rtol=1e-9)


def test_openmp_matches_serial():
    g = bc_amd.gen_rmat(10, 8 << 10, seed=11)
    src = bc_amd.sample_sources(g.n, 40, 3)
    a = O.brandes(g, src, threads=1)
    b = O.brandes(g, src, threads=4)
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-9)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and a[3] == b[3]
