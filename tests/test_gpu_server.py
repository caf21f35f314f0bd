"""Serving on the GPU: every request's result equals the same request served
alone (bit-identical — each request's kernels run in order on its own stream
and workspace, whatever runs beside it) and the oracle, with several
requests overlapping on the device."""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


def served_alone(graphs, j, kind, src):
    """The same request on a fresh one-slot server, nothing else running."""
    srv = bc_amd.Server(slots=1, hbm_budget=64 << 30)
    gids = [srv.add_graph(g) for g in graphs]
    rid = srv.submit(gids[j], kind, src)
    srv.wait(rid)
    st, bc, far, _ = srv.result(rid)
    srv.close()
    assert st == 0
    return bc, far


def test_results_same_alone_and_in_mix(ctx):
    graphs = [bc_amd.gen_rmat(12, 16 << 12, seed=21), bc_amd.gen_rmat(14, 16 << 14, seed=22),
              bc_amd.gen_blocktree(120, 1.5, 8, 3000, 6.0, 0.3, 0.5, seed=23),
              bc_amd.gen_grid(40, 50, 0.05, 3, seed=24)]
    srv = bc_amd.Server(slots=4, hbm_budget=32 << 30)
    gids = [srv.add_graph(g) for g in graphs]
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(24):
        j = int(rng.integers(0, len(graphs)))
        g = graphs[j]
        if i % 2 == 0:
            src = bc_amd.sample_sources(g.n, int(rng.integers(1, 200)), 100 + i)
            reqs.append((srv.submit(gids[j], bc_amd.CENT_Q_BC, src), j, "bc", src))
        elif i % 4 == 1:
            reqs.append((srv.submit(gids[j], bc_amd.CENT_Q_CC, None), j, "cc_all", None))
        else:
            src = bc_amd.sample_sources(g.n, int(rng.integers(1, 100)), 200 + i)
            reqs.append((srv.submit(gids[j], bc_amd.CENT_Q_CC, src), j, "cc", src))
    print("[gpu] server: 24 requests submitted", flush=True)
    dgs = [ctx.upload(g) for g in graphs]
    plain_far = {}
    for rid, j, kind, src in reqs:
        info = srv.wait(rid)
        assert info["state"] == "done", info
        st, bc, far, info = srv.result(rid)
        assert st == 0
        g = graphs[j]
        if kind == "bc":
            abc, afar = served_alone(graphs, j, bc_amd.CENT_Q_BC, src)
            assert np.array_equal(bc, abc) and np.array_equal(far, afar)
            obc, ofar, _, _ = O.brandes(g, src, threads=8)
            np.testing.assert_allclose(bc, obc, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(obc).max()))
            assert np.array_equal(far, ofar)
        elif kind == "cc":
            assert np.array_equal(far, ctx.cc(dgs[j], src)[0])
            assert np.array_equal(far, O.farness(g, src))
        else:
            if j not in plain_far:
                plain_far[j] = O.weighted_farness(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64))
            assert np.array_equal(far, plain_far[j])
    st = srv.stats()
    print(f"[gpu] server stats {st}", flush=True)
    assert st["completed"] == 24 and st["max_running"] >= 2
    srv.close()


@pytest.mark.timeout(600)
def test_mix_with_pool_sized_graphs(ctx):
    """Requests on graphs of serving-pool size (>= 1e5 vertices: an R-MAT with
    16 draws per vertex and a 4000-block tree) mixed with small ones.  Every
    result equals the same call alone; BC and source farness equal the
    oracle's; all-vertex farness equals the plain all-source engine call and,
    on a sample of 1024 vertices, the oracle's BFS."""
    graphs = [bc_amd.gen_rmat(18, 16 << 18, seed=31),
              bc_amd.gen_blocktree(4000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=32),
              bc_amd.gen_rmat(12, 16 << 12, seed=33)]
    assert graphs[0].n >= 100_000 and graphs[1].n >= 100_000, [g.n for g in graphs]
    print(f"[gpu] server pool-size mix: n = {[g.n for g in graphs]}", flush=True)
    srv = bc_amd.Server(slots=4, hbm_budget=64 << 30)
    gids = [srv.add_graph(g) for g in graphs]
    plan = [(0, "bc", 512), (1, "cc_all", 0), (0, "cc_all", 0), (2, "bc", 300), (1, "bc", 256),
            (0, "bc", 64), (2, "cc_all", 0), (1, "cc", 700), (0, "cc", 1024), (1, "cc_all", 0)]
    reqs = []
    for i, (j, kind, k) in enumerate(plan):
        g = graphs[j]
        src = bc_amd.sample_sources(g.n, k, 300 + i) if k else None
        qk = bc_amd.CENT_Q_BC if kind == "bc" else bc_amd.CENT_Q_CC
        reqs.append((srv.submit(gids[j], qk, src), j, kind, src))
    dgs = [ctx.upload(g) for g in graphs]
    all_far = {}
    for rid, j, kind, src in reqs:
        info = srv.wait(rid)
        assert info["state"] == "done", info
        st, bc, far, info = srv.result(rid)
        assert st == 0
        g = graphs[j]
        if kind == "bc":
            abc, afar = served_alone(graphs, j, bc_amd.CENT_Q_BC, src)
            assert np.array_equal(bc, abc) and np.array_equal(far, afar)
            obc, ofar, _, _ = O.brandes(g, src, threads=16)
            np.testing.assert_allclose(bc, obc, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(obc).max()))
            assert np.array_equal(far, ofar)
        elif kind == "cc":
            assert np.array_equal(far, O.weighted_farness(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64), src))
        else:
            if j not in all_far:
                all_far[j] = ctx.cc(dgs[j], None)[0]
                sample = bc_amd.sample_sources(g.n, 1024, 77)
                want = O.weighted_farness(g.n, g.row_ptr, g.col, np.ones(g.n, np.int64), sample)
                assert np.array_equal(all_far[j][sample], want)
            assert np.array_equal(far, all_far[j])
        print(f"[gpu] server pool-size mix: {kind} on graph {j} ok", flush=True)
    st = srv.stats()
    assert st["completed"] == len(plan) and st["max_running"] >= 2, st
    srv.close()
