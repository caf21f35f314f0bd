"""Server admission and scheduling on the CPU (dry-run mode: requests sleep for
a modelled time; no device is touched and no results are produced)."""
import time

import numpy as np
import pytest

import bc_amd
from conftest import probe_depth


def small_graph(n=200, seed=0):
    return bc_amd.gen_rmat(8, 8 << 8, seed=seed)


def test_concurrency_bounded_by_slots():
    s = bc_amd.Server(slots=3, hbm_budget=1 << 40, dryrun=True)
    gid = s.add_graph(small_graph())
    rids = [s.submit(gid, bc_amd.CENT_Q_BC, [0, 1, 2]) for _ in range(12)]
    infos = [s.wait(r) for r in rids]
    st = s.stats()
    assert st["completed"] == 12 and st["max_running"] <= 3 and st["running"] == 0
    assert all(i["state"] == "done" for i in infos)
    # FIFO admission: start times follow submission order
    starts = [i["t_start_ms"] for i in infos]
    assert starts == sorted(starts)
    for r in rids:
        status, bc, far, info = s.result(r)
        assert status == 0
    s.close()


def test_budget_limits_admission_and_oversized_fails():
    g = small_graph()
    src = [0, 1, 2]
    # closeness from listed sources runs the plain call: its reservation is
    # the engine's estimate for it (max_slots 256, one stream)
    s = bc_amd.Server(slots=4, hbm_budget=1 << 40, dryrun=True)
    gid = s.add_graph(g)
    F = s.wait(s.submit(gid, bc_amd.CENT_Q_CC, src))["footprint"]
    graph_bytes = s.stats()["used_graph"]
    assert F == bc_amd.workspace_estimate(g, bc_amd.CENT_TUNE_CC, len(src), probe_depth(g), max_slots=256,
                                          one_stream=True) > 0
    # betweenness goes through the reduction, whose GPU build reserves its scratch first
    first = s.wait(s.submit(gid, bc_amd.CENT_Q_BC, [0]))["footprint"]
    assert first >= bc_amd.reduce_gpu_workspace_estimate(g)
    s.close()
    # a budget that fits one reservation at a time serialises the requests
    s = bc_amd.Server(slots=4, hbm_budget=graph_bytes + F + F // 2, dryrun=True)
    gid = s.add_graph(g)
    rids = [s.submit(gid, bc_amd.CENT_Q_CC, src) for _ in range(6)]
    for r in rids:
        assert s.wait(r)["state"] == "done"
    st = s.stats()
    assert st["max_running"] == 1 and st["used_workspace"] == 0
    s.close()
    # one that never fits it fails the request instead of blocking the queue
    s = bc_amd.Server(slots=2, hbm_budget=graph_bytes + F - 1, dryrun=True)
    gid = s.add_graph(g)
    r = s.submit(gid, bc_amd.CENT_Q_CC, src)
    info = s.wait(r)
    assert info["state"] == "failed"
    status, _, _, _ = s.result(r)
    assert status == bc_amd.CENT_E_OOM
    s.close()


def test_queue_full_and_cancel():
    s = bc_amd.Server(slots=1, hbm_budget=1 << 40, max_queue=2, dryrun=True)
    big = bc_amd.gen_rmat(12, 16 << 12, seed=1)
    gid = s.add_graph(big)
    first = s.submit(gid, bc_amd.CENT_Q_BC, None)   # all sources: ~0.3 s modelled
    time.sleep(0.05)
    q1 = s.submit(gid, bc_amd.CENT_Q_BC, [1])
    q2 = s.submit(gid, bc_amd.CENT_Q_BC, [2])
    with pytest.raises(bc_amd.CentError) as e:
        s.submit(gid, bc_amd.CENT_Q_BC, [3])
    assert e.value.status == bc_amd.CENT_E_BUSY
    s.cancel(q1)
    assert s.wait(q1)["state"] == "cancelled"
    assert s.wait(q2)["state"] == "done" and s.wait(first)["state"] == "done"
    st = s.stats()
    assert st["rejected"] == 1 and st["cancelled"] == 1 and st["completed"] == 2
    assert s.result(q1)[0] == bc_amd.CENT_E_CANCELLED
    with pytest.raises(bc_amd.CentError):
        s.submit(99, bc_amd.CENT_Q_BC, [0])
    s.close()


def test_graph_cache_hits_and_evictions():
    gs = [small_graph(seed=i) for i in range(3)]
    src = [0, 1, 2]
    s = bc_amd.Server(slots=1, hbm_budget=1 << 40, dryrun=True)
    sizes, fp = [], []
    for g in gs:
        fp.append(s.wait(s.submit(s.add_graph(g), bc_amd.CENT_Q_CC, src))["footprint"])
        sizes.append(s.stats()["used_graph"] - sum(sizes))
    s.close()
    # room for the largest closeness workspace plus about one cached graph
    s = bc_amd.Server(slots=1, hbm_budget=max(fp) + max(sizes) * 3 // 2, dryrun=True)
    gids = [s.add_graph(g) for g in gs]
    for gid in [gids[0], gids[0], gids[1], gids[0], gids[2]]:
        r = s.submit(gid, bc_amd.CENT_Q_CC, src)
        assert s.wait(r)["state"] == "done"
    st = s.stats()
    assert st["graph_hits"] >= 1 and st["graph_evictions"] >= 1
    assert st["graph_uploads"] + st["graph_hits"] == 5
    s.close()
