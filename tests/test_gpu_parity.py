"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle.

Bar: farness / distance sums / reached counts bit-exact (int64); BC within
1e-9 relative (FP64, summation order differs from the oracle's).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def assert_bc_close(got, want, what=""):
    scale = max(1.0, float(np.abs(want).max()) if len(want) else 1.0)
    np.testing.assert_allclose(got, want, rtol=BC_RTOL, atol=BC_RTOL * scale, err_msg=what)


FIXTURES = ["karate.txt", "dolphins.txt", "football.txt", "polbooks.txt", "testdata.txt",
            "testdata1.txt", "testdata2.txt", "ia-fb-messages.mtx"]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_all_sources(ctx, name):
    print(f"[gpu] fixture {name}", flush=True)
    g = bc_amd.buildGraph(O.dataset_path(name))
    dg = ctx.upload(g)
    src = g.sources()
    for mode in (bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT):
        bc, far, ds = ctx.bc(dg, src, mode=mode, dist_sum=True)
        obc, ofar, ods, _ = O.brandes(g, src, mode=mode)
        assert_bc_close(bc, obc, f"{name} mode {mode}")
        assert np.array_equal(far, ofar)
        assert np.array_equal(ds, ods)
    cfar, cds = ctx.cc(dg, src, dist_sum=True)
    assert np.array_equal(cfar, ofar)
    assert np.array_equal(cds, ods)


def test_reference_answers_on_gpu(ctx):
    """karate_BC_Score.txt (reference formula, FP32 print) and CCs[7] = 86."""
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    BCs = np.zeros(g.n)
    bc_amd.computeBC(g, BCs, compat=True)
    lines = open(os.path.join(GOLDEN, "karate_BC_Score.txt")).read().split()
    ref = np.array([float(x) for x in lines[2:]])
    np.testing.assert_allclose(BCs[1:35], ref, rtol=2e-6, atol=2e-6)
    CCs = np.zeros(g.n, np.int64)
    bc_amd.computeCC(g, CCs)
    assert CCs[7] == 86
    ref = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))["karate.txt"]
    assert CCs[1:35].tolist() == ref["cc"]


def test_musae_git_known_answer(ctx):
    g = bc_amd.buildGraph(O.dataset_path("musae_git.txt"))
    dg = ctx.upload(g)
    src = np.array([7, 0, 1, 37699], np.int32)
    far, _ = ctx.cc(dg, src)
    assert far[0] == 117726
    assert np.array_equal(far, O.farness(g, src))


def test_rmat14_all_sources(ctx):
    """R-MAT 2^14, avg degree 8, LCC, all sources: exact BC + farness."""
    g = bc_amd.gen_rmat(14, 8 << 14, seed=1)
    print(f"[gpu] rmat14 n={g.n} m={g.m}", flush=True)
    dg = ctx.upload(g)
    bc, far, ds = ctx.bc(dg, None, dist_sum=True)
    obc, ofar, ods, _ = O.brandes(g, None, threads=min(16, O.max_threads()))
    assert_bc_close(bc, obc, "rmat14")
    assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
    cfar, cds = ctx.cc(dg, None, dist_sum=True)
    assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)


def test_sampled_ragged_batches(ctx):
    g = bc_amd.gen_rmat(16, 16 << 16, seed=5)
    dg = ctx.upload(g)
    for k in (1, 63, 64, 65, 200):
        src = bc_amd.sample_sources(g.n, k, 100 + k)
        bc, far, ds = ctx.bc(dg, src, dist_sum=True)
        obc, ofar, ods, _ = O.brandes(g, src, threads=min(16, O.max_threads()))
        assert_bc_close(bc, obc, f"k={k}")
        assert np.array_equal(far, ofar) and np.array_equal(ds, ods)


def test_edge_cases(ctx):
    # disconnected graph with an isolated vertex; duplicate sources; empty source list
    g = bc_amd.CSR.from_edges(9, [0, 1, 3, 4, 4, 6], [1, 2, 4, 5, 6, 7])
    dg = ctx.upload(g)
    src = np.array([8, 0, 0, 4, 7, 2], np.int32)
    bc, far, ds = ctx.bc(dg, src, dist_sum=True)
    obc, ofar, ods, _ = O.brandes(g, src)
    assert_bc_close(bc, obc)
    assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
    assert far[0] == 0
    bc, far, ds = ctx.bc(dg, np.zeros(0, np.int32), dist_sum=True)
    assert np.all(bc == 0) and len(far) == 0 and np.all(ds == 0)
    with pytest.raises(bc_amd.CentError) as e:
        ctx.bc(dg, np.array([9], np.int32))
    assert e.value.status == bc_amd.CENT_E_INVAL
    # BC accumulates into the caller's array (Brandes.c:224 contract)
    acc = np.full(g.n, 1.5)
    ctx.bc(dg, src, bc=acc)
    assert_bc_close(acc - 1.5, obc)


def test_deterministic(ctx):
    g = bc_amd.gen_rmat(15, 16 << 15, seed=9)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 130, 1)
    a = ctx.bc(dg, src)[0]
    b = ctx.bc(dg, src)[0]
    assert np.array_equal(a, b)


def test_large_rmat_properties(ctx):
    """R-MAT 2^20 (the sampled-BC workload's graph) — size-independent checks:
    sum identity per source set, and a 64-source subset bit-exact vs the oracle."""
    g = bc_amd.gen_rmat(20, 16 << 20, seed=2)
    print(f"[gpu] rmat20 n={g.n} m={g.m}", flush=True)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 256, 2)
    ctx.reset_stats()
    bc, far, ds = ctx.bc(dg, src, dist_sum=True)
    st = ctx.stats()
    reached = st["reached"]
    # sum_v delta_s(v) = sum_{t != s} (d(s,t) - 1) = farness(s) - (reached_s - 1)
    assert np.isclose(bc.sum(), far.sum() - (reached - len(src)), rtol=1e-9)
    assert ds.sum() == far.sum()
    sub = src[:64]
    obc, ofar, ods, odag = O.brandes(g, sub, threads=min(16, O.max_threads()))
    bc2, far2, ds2 = ctx.bc(dg, sub, dist_sum=True)
    assert_bc_close(bc2, obc, "rmat20 subset")
    assert np.array_equal(far2, ofar) and np.array_equal(ds2, ods)


@pytest.mark.parametrize("graph", ["rmat", "grid"])
@pytest.mark.parametrize("order", ["keep", "degree", "bfs"])
def test_device_order_does_not_change_results(ctx, graph, order):
    """Default device layout (degree order for the shallow R-MAT, BFS order for
    the deep grid) vs each forced layout, on the step the graph runs by
    default (dense 64-lane for R-MAT, source-parallel for the grid), BC and
    closeness, against the oracle."""
    print(f"[gpu] device order {order} on {graph}", flush=True)
    if graph == "rmat":
        g = bc_amd.gen_rmat(15, 16 << 15, seed=4)
    else:
        g = bc_amd.gen_grid(256, 256, 0.05, 20, seed=6)
    src = bc_amd.sample_sources(g.n, 100, 4)
    a = ctx.bc(ctx.upload(g), src, dist_sum=True)
    dg = ctx.upload(g, order=order)
    b = ctx.bc(dg, src, dist_sum=True)
    assert_bc_close(a[0], b[0])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    cf, cd = ctx.cc(dg, src, dist_sum=True)
    assert np.array_equal(cf, a[1]) and np.array_equal(cd, a[2])
    obc, ofar, ods, _ = O.brandes(g, src)
    assert_bc_close(a[0], obc)
    assert np.array_equal(a[1], ofar) and np.array_equal(a[2], ods)


def test_device_order_flags_exclusive(ctx):
    g = bc_amd.gen_rmat(8, 8 << 8, seed=1)
    h = C.c_void_p()
    flags = bc_amd.CENT_GRAPH_DEGREE_ORDER | bc_amd.CENT_GRAPH_BFS_ORDER
    rc = bc_amd._lib.cent_graph_create(ctx._h, bc_amd._ptr(g.row_ptr, C.c_int32), bc_amd._ptr(g.col, C.c_int32),
                                       g.n, g.nnz, flags, C.byref(h))
    assert rc == bc_amd.CENT_E_INVAL and not h.value
    with pytest.raises(ValueError):
        ctx.upload(g, order="random")


@pytest.mark.parametrize("frontier", ["dense", "sparse", "source"])
def test_frontier_modes_agree(ctx, frontier):
    """Dense pull levels, frontier-queue (hipGraph-replayed) levels and
    source-parallel workgroups give the oracle's results on a power-law graph,
    a high-diameter grid, a disconnected graph with duplicate sources, and in
    the reference-compat formula."""
    cases = [("rmat13", bc_amd.gen_rmat(13, 16 << 13, seed=3), 150),
             ("grid", bc_amd.gen_grid(64, 96, 0.05, 3, seed=4), 130),
             ("path", bc_amd.CSR.from_edges(300, np.arange(299), np.arange(1, 300)), 70)]
    for name, g, k in cases:
        print(f"[gpu] frontier={frontier} {name} n={g.n}", flush=True)
        dg = ctx.upload(g)
        src = bc_amd.sample_sources(g.n, min(k, g.n), 7)
        src = np.concatenate([src, src[:3]])  # duplicate sources in the last batch
        for mode in (bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT):
            bc, far, ds = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier=frontier)
            obc, ofar, ods, _ = O.brandes(g, src, mode=mode, threads=min(16, O.max_threads()))
            assert_bc_close(bc, obc, f"{name} {frontier} mode {mode}")
            assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
        cfar, cds = ctx.cc(dg, src, dist_sum=True, frontier=frontier)
        assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)
    g = bc_amd.CSR.from_edges(9, [0, 1, 3, 4, 4, 6], [1, 2, 4, 5, 6, 7])
    src = np.array([8, 0, 0, 4, 7, 2], np.int32)
    bc, far, ds = ctx.bc(ctx.upload(g), src, dist_sum=True, frontier=frontier)
    obc, ofar, ods, _ = O.brandes(g, src)
    assert_bc_close(bc, obc)
    assert np.array_equal(far, ofar) and np.array_equal(ds, ods)


@pytest.mark.parametrize("slots", [1, 3, 7])
def test_source_mode_slot_reuse(ctx, slots):
    """Source-parallel slots run many sources each (restoring their arrays in
    between); rows longer than the wavefront threshold take the long-row path
    (R-MAT hubs).  Results match the oracle for every slot count."""
    for name, g, k in [("rmat12", bc_amd.gen_rmat(12, 16 << 12, seed=5), 90),
                       ("grid", bc_amd.gen_grid(40, 50, 0.05, 5, seed=6), 75)]:
        print(f"[gpu] source slots={slots} {name} n={g.n}", flush=True)
        dg = ctx.upload(g)
        src = bc_amd.sample_sources(g.n, k, 9)
        for mode in (bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT):
            bc, far, ds = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier="source", max_slots=slots)
            obc, ofar, ods, _ = O.brandes(g, src, mode=mode, threads=min(16, O.max_threads()))
            assert_bc_close(bc, obc, f"{name} slots {slots} mode {mode}")
            assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
        cfar, cds = ctx.cc(dg, src, dist_sum=True, frontier="source", max_slots=slots)
        assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)


def test_source_mode_closeness_all_sources(ctx):
    """Closeness-only source-parallel step at its widest: every vertex of a
    small graph is a source, so the slot rule runs 1024 slots with their bit
    maps side by side (map words shared by up to 32 vertices); isolated
    vertices and a second component included.  Farness and distance sums
    bit-exact against the oracle, and against the 64-lane step."""
    g0 = bc_amd.gen_rmat(12, 16 << 12, seed=8, lcc=False)
    n = g0.n + 40  # 40 more vertices: a path of 37 (its own component) and 3 isolated
    u = np.concatenate([np.repeat(np.arange(g0.n), np.diff(g0.row_ptr)), np.arange(g0.n, g0.n + 36)])
    v = np.concatenate([g0.col, np.arange(g0.n + 1, g0.n + 37)])
    keep = u < v
    g = bc_amd.CSR.from_edges(n, u[keep], v[keep])
    print(f"[gpu] source-mode closeness, all {g.n} sources", flush=True)
    dg = ctx.upload(g)
    far, ds = ctx.cc(dg, None, dist_sum=True, frontier="source")
    ofar = O.farness(g, np.arange(g.n, dtype=np.int32))
    assert np.array_equal(far, ofar)
    assert ds.sum() == far.sum()
    f2, d2 = ctx.cc(dg, None, dist_sum=True, frontier="dense")
    assert np.array_equal(f2, far) and np.array_equal(d2, ds)


def test_closeness_wide_groups(ctx):
    """The interleaved closeness step at its widest launch groups: all
    sources of an R-MAT 2^14 graph (groups of 32 batches) and 16437 sources
    with repeats (groups of 64 batches and a ragged last one), farness and
    distance sums bit-exact against the source-parallel step, and a subset
    against the oracle."""
    g = bc_amd.gen_rmat(14, 16 << 14, seed=12)
    dg = ctx.upload(g)
    rng = np.random.default_rng(3)
    for src in (None, rng.integers(0, g.n, 16437).astype(np.int32)):
        k = g.n if src is None else len(src)
        print(f"[gpu] closeness wide groups n={g.n} k={k}", flush=True)
        far, ds = ctx.cc(dg, src, dist_sum=True, frontier="dense")
        far2, ds2 = ctx.cc(dg, src, dist_sum=True, frontier="source")
        assert np.array_equal(far, far2) and np.array_equal(ds, ds2)
        sub = np.arange(g.n, dtype=np.int32)[:97] if src is None else src[-97:]
        want = O.farness(g, sub)
        assert np.array_equal(far[:97] if src is None else far[-97:], want)


def test_source_mode_deterministic(ctx):
    """Same slot count, same bits: BC partials are reduced in slot order, and
    every row longer than the wavefront threshold is summed by a wavefront,
    also on levels with more such rows than the LDS list holds (R-MAT 2^14:
    over a thousand hub rows on a level)."""
    for g, slots in ((bc_amd.gen_grid(48, 48, 0.05, 8, seed=2), 16), (bc_amd.gen_rmat(14, 16 << 14, seed=22), 256)):
        print(f"[gpu] source-mode determinism n={g.n}", flush=True)
        dg = ctx.upload(g)
        src = bc_amd.sample_sources(g.n, 200, 3)
        a = ctx.bc(dg, src, frontier="source", max_slots=slots)[0]
        b = ctx.bc(dg, src, frontier="source", max_slots=slots)[0]
        assert np.array_equal(a, b)


def test_source_mode_deep_closeness(ctx):
    """A BFS deeper than the betweenness step's level table (path of 70000
    vertices from an end): the closeness step keeps no level table and
    finishes it itself: farness n(n-1)/2, distance sums 0 .. n-1."""
    n = 70000
    g = bc_amd.CSR.from_edges(n, np.arange(n - 1), np.arange(1, n))
    dg = ctx.upload(g, keep_order=True)
    print("[gpu] source-mode deep closeness, path 70000", flush=True)
    far, ds = ctx.cc(dg, [0], dist_sum=True, frontier="source")
    assert far[0] == n * (n - 1) // 2
    assert np.array_equal(ds, np.arange(n, dtype=np.int64))


def diamond_chain(k):
    """k diamonds h_i - {a_i, b_i} - h_{i+1} in a row: 2^i shortest paths from
    h_0 to h_i, so FP64 path counts leave their range after 1023 diamonds."""
    u, v = [], []
    for i in range(k):
        h, a, b, h2 = 3 * i, 3 * i + 1, 3 * i + 2, 3 * i + 3
        u += [h, h, a, b]
        v += [a, b, h2, h2]
    return bc_amd.CSR.from_edges(3 * k + 1, np.array(u), np.array(v))


@pytest.mark.parametrize("frontier", ["dense", "sparse", "source"])
def test_path_counts_beyond_fp64_are_scaled(ctx, frontier):
    """2^1100 shortest paths from h_0 to h_1100: the counts are scaled per
    level, not refused.  From h_0 alone the dependencies have a closed form:
    delta(h_i) = 3 (k - i) vertices lie beyond h_i, and each middle vertex of
    diamond i carries half of the paths into h_{i+1}:
    delta(a_i) = (1 + 3 (k - i - 1)) / 2."""
    print(f"[gpu] scaled counts frontier={frontier}", flush=True)
    k = 1100
    g = diamond_chain(k)
    dg = ctx.upload(g, keep_order=True)
    bc, far, _ = ctx.bc(dg, [0], frontier=frontier)
    want = np.zeros(g.n)
    for i in range(k):
        want[3 * i + 1] = want[3 * i + 2] = (1 + 3 * (k - i - 1)) / 2
        if i > 0:
            want[3 * i] = 3 * (k - i)
    assert np.all(np.isfinite(bc))
    np.testing.assert_allclose(bc, want, rtol=1e-12, atol=0)
    # farness in closed form: h_i at 2i, the middle vertices of diamond i at 2i + 1
    assert far[0] == sum(2 * i for i in range(1, k + 1)) + sum(2 * (2 * i + 1) for i in range(k))
    cfar, _ = ctx.cc(dg, [0], frontier=frontier)
    assert cfar[0] == far[0]
    # below 2^1024 the counts are exact powers of two and BC equals the oracle's
    g2 = diamond_chain(1000)
    dg2 = ctx.upload(g2, keep_order=True)
    bc, far2, _ = ctx.bc(dg2, [0], frontier=frontier)
    obc, ofar, _, _ = O.brandes(g2, [0])
    assert_bc_close(bc, obc)
    assert np.array_equal(far2, ofar)


@pytest.mark.parametrize("frontier", ["dense", "sparse"])
def test_batch_grouping_keeps_caller_order(ctx, frontier):
    """The 64-lane steps regroup sources into batches; farness comes back in
    the caller's order and every result matches the given-order run."""
    g = bc_amd.gen_rmat(13, 16 << 13, seed=8)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 300, 12)
    a = ctx.bc(dg, src, dist_sum=True, frontier=frontier)
    b = ctx.bc(dg, src, dist_sum=True, frontier=frontier, keep_source_order=True)
    obc, ofar, ods, _ = O.brandes(g, src)
    for bc, far, ds in (a, b):
        assert_bc_close(bc, obc)
        assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
    c = ctx.cc(dg, src, frontier=frontier)[0]
    assert np.array_equal(c, ofar)


@pytest.mark.parametrize("what", ["bc", "cc"])
def test_second_stream_matches_one_stream(ctx, what):
    """Launch groups alternating over two streams (default) against one
    stream: farness and distance sums exact, BC equal up to the summation
    order of the two streams' partial sums, and the two-stream result the
    same bits on every run."""
    g = bc_amd.gen_rmat(13, 16 << 13, seed=31)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 300 if what == "bc" else 1100, 9)
    if what == "bc":  # the 64-lane step (the default for betweenness on graphs this small is per source)
        a = ctx.bc(dg, src, dist_sum=True, frontier="dense")
        b = ctx.bc(dg, src, dist_sum=True, one_stream=True, frontier="dense")
        c = ctx.bc(dg, src, dist_sum=True, frontier="dense")
        np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=1e-12 * max(1.0, np.abs(b[0]).max()))
        assert np.array_equal(a[0], c[0])
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        obc, ofar, ods, _ = O.brandes(g, src)
        assert_bc_close(a[0], obc)
        assert np.array_equal(a[1], ofar) and np.array_equal(a[2], ods)
    else:
        a = ctx.cc(dg, src, dist_sum=True)
        b = ctx.cc(dg, src, dist_sum=True, one_stream=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(a[0], O.farness(g, src))


def test_grid2048_source_mode_properties(ctx):
    """The high-diameter workload's graph (2048 x 2048 grid, 5% of edges removed,
    1000 shortcuts) in the source-parallel step: the per-source sum identity
    and distance-sum total at full size, and an 8-source subset against the
    oracle and against the dense step."""
    g = bc_amd.gen_grid(2048, 2048, 0.05, 1000, seed=4)
    print(f"[gpu] grid2048 n={g.n} source mode", flush=True)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 128, 4)
    ctx.reset_stats()
    bc, far, ds = ctx.bc(dg, src, dist_sum=True, frontier="source")
    reached = ctx.stats()["reached"]
    assert np.isclose(bc.sum(), far.sum() - (reached - len(src)), rtol=1e-9)
    assert ds.sum() == far.sum()
    sub = src[:8]
    obc, ofar, ods, _ = O.brandes(g, sub, threads=min(16, O.max_threads()))
    bc2, far2, ds2 = ctx.bc(dg, sub, dist_sum=True, frontier="source")
    assert_bc_close(bc2, obc, "grid2048 subset")
    assert np.array_equal(far2, ofar) and np.array_equal(ds2, ods)
    bc3, far3, _ = ctx.bc(dg, sub, frontier="dense")
    assert_bc_close(bc3, bc2, "grid2048 dense vs source")
    assert np.array_equal(far3, far2)


def test_grid512_source_mode_matches_oracle(ctx):
    g = bc_amd.gen_grid(512, 512, 0.05, 50, seed=9)
    print(f"[gpu] grid512 n={g.n} source mode", flush=True)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 300, 9)
    for mode in (bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT):
        bc, far, ds = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier="source")
        obc, ofar, ods, _ = O.brandes(g, src, mode=mode, threads=min(16, O.max_threads()))
        assert_bc_close(bc, obc, f"grid512 mode {mode}")
        assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
