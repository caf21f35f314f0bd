#!/usr/bin/env python3
"""Benchmark: sampled Brandes BC + closeness on R-MAT 2^20 (avg degree 16).

Workload (one step): 4096 uniformly sampled sources (seed 2) on an R-MAT graph
with 2^20 ids and 16 x 2^20 edge draws (a=0.57, b=c=0.19, d=0.05, seed 2),
self-loops and duplicates removed, symmetrised, largest component kept, int32
CSR.  One step computes, for every source, its Brandes dependencies
(accumulated into FP64 BC) and its farness (int64) plus the per-vertex int64
distance sums, all in one fused multi-source pass on the GPU.  The graph is
resident in HBM before the timed region and the step's outputs stay there
(cent_bc_device); the same step with host-array outputs (cent_bc, scores
copied back over PCIe) is reported beside it as pcie_inclusive.

Prints one JSON line.  value = TEPS = sources x m / seconds per step, with m the
undirected edges of the (connected) graph.  roofline = the dominant level
pass's HBM-side bytes per launch (rocprofv3 counters of this workload,
profiles/rmat20_traffic.json, used only while its key matches the tree) over
its average launch duration in a one-stream step (HIP events); the per-source
bytes model (forward 16m + 20n, backward 16m + 40n + 16 per shortest-path
DAG edge) is reported beside it as an effective rate only.
cpu_baseline = the CPU oracle (a C restatement of the same Brandes pass)
timed on a bounded source sample on this host's cores, 16 threads and one.
serving = the request-server workload (Poisson stream, p50/p99, served vs
plain).  Any result that differs from its check makes the exit status 1.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


def timed(fn, reps: int):
    fn()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return sum(ts) / len(ts)


def reduced_work_edges(red) -> int:
    """Source-edge pairs the reduced traversal walks: every packed vertex of a
    block is a BFS source over that block's edges, so sum over blocks of
    n_b * m_b ("work TEPS" counts these; "effective TEPS" counts n * m of
    the original graph)."""
    import numpy as np
    b = red.boff.astype(np.int64)
    rp = red.prp.astype(np.int64)
    return int((np.diff(b) * ((rp[b[1:]] - rp[b[:-1]]) // 2)).sum())


def extra_blocktree(ctx, args):
    """Block-tree workload: 2000 blocks (Pareto(1.5) sizes in [8, 50000], ER avg
    degree 6 + Hamiltonian cycle), random tree of shared articulation vertices,
    +30% pendant chains (Geometric(0.5) lengths), seed 3; all-source closeness
    through D1 folding + block split.  Compares one persistent launch over the
    work-item queue with level-synchronous launches per work item."""
    import numpy as np
    import bc_amd
    t0 = time.time()
    g = bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3)
    red = bc_amd.Reduced(g)
    log(f"blocktree: n={g.n} m={g.m} blocks={red.nblocks} aps={red.n_ap} folded={red.nfold} "
        f"packed={red.np} items={red.nitems} (built in {time.time() - t0:.1f}s)")
    res = {}
    far_p = ctx.cc_reduced(red, bc_amd.CENT_SCHED_PERSISTENT)
    far_i = ctx.cc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM)
    assert (far_p == far_i).all()
    far_s = ctx.cc_reduced(red, bc_amd.CENT_SCHED_SHARED)
    assert (far_s == far_p).all()
    tp = timed(lambda: ctx.cc_reduced(red, bc_amd.CENT_SCHED_PERSISTENT), 5)
    ti = timed(lambda: ctx.cc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM), 2)
    tsh = timed(lambda: ctx.cc_reduced(red, bc_amd.CENT_SCHED_SHARED), 5)
    dg = ctx.upload(g)
    tplain = timed(lambda: ctx.cc(dg, None), 1)
    log(f"blocktree: persistent {tp * 1e3:.2f} ms, per-item {ti * 1e3:.1f} ms, "
        f"neighbour-sharing {tsh * 1e3:.2f} ms, plain all-source MS-BFS {tplain * 1e3:.1f} ms")
    # BC + closeness of every vertex through the reduction
    bp, fp = ctx.bc_reduced(red, bc_amd.CENT_SCHED_PERSISTENT)
    bi, fi = ctx.bc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM)
    assert (fp == far_p).all() and (fi == far_p).all()
    import numpy as np
    assert np.allclose(bp, bi, rtol=1e-9, atol=1e-9 * max(1.0, float(np.abs(bp).max())))
    tbp = timed(lambda: ctx.bc_reduced(red, bc_amd.CENT_SCHED_PERSISTENT), 3)
    tbi = timed(lambda: ctx.bc_reduced(red, bc_amd.CENT_SCHED_PER_ITEM), 1)
    log(f"blocktree BC+CC: persistent {tbp * 1e3:.2f} ms, per-item {tbi * 1e3:.1f} ms")
    n, m = g.n, g.m
    work = reduced_work_edges(red)
    # the reference's own compute_D1_AP_CC_shareBased_DegreeOrder (Seq_CC/CC.cpp:1728-2052),
    # built from its sources into oracle/_ref, on the same graph, one host thread
    ref_cc = os.path.join(ROOT, "oracle", "_ref", "ref_cc")
    if os.path.exists(ref_cc):
        import subprocess
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "blocktree.txt")
            bc_amd.write_edgelist(path, g)
            t0 = time.perf_counter()
            subprocess.run([ref_cc, "d1ap", path, os.path.join(td, "d1ap.txt")], check=True, timeout=900,
                           stdout=subprocess.DEVNULL)
            t_ref = time.perf_counter() - t0
            rcc = np.loadtxt(os.path.join(td, "d1ap.txt"), dtype=np.int64)
        res["cpu_baseline"] = {
            "kind": "reference", "cores": 1, "s": round(t_ref, 3),
            "sample": "whole workload: the reference's compute_D1_AP_CC_shareBased_DegreeOrder "
                      f"(built from its sources), CPU {cpu_model()}",
            "farness_equal": bool(np.array_equal(rcc[:, 1], far_p[rcc[:, 0]])),
            "speedup_persistent": round(t_ref / tp, 1)}
        log(f"blocktree: reference D1/AP closeness {t_ref:.2f} s on one core, "
            f"farness equal: {res['cpu_baseline']['farness_equal']}")
    res.update({
        "workload": "all-source closeness, block tree 2000 blocks (D1 + AP reduced), seed 3",
        "n": n, "m": m, "blocks": red.nblocks, "articulation": red.n_ap, "folded": red.nfold,
        "packed_vertices": red.np, "items": red.nitems,
        "persistent_ms": round(tp * 1e3, 3), "per_item_ms": round(ti * 1e3, 3),
        "neighbour_sharing_ms": round(tsh * 1e3, 3),
        "plain_unreduced_ms": round(tplain * 1e3, 3),
        "queries_per_s_persistent": round(1.0 / tp, 2), "queries_per_s_per_item": round(1.0 / ti, 2),
        "speedup_vs_per_item": round(ti / tp, 2),
        "effective_TEPS_persistent": round(n * m / tp, 1),
        "reduced_work_edges": work, "work_TEPS_persistent": round(work / tp, 1),
        "bc_cc_persistent_ms": round(tbp * 1e3, 3), "bc_cc_per_item_ms": round(tbi * 1e3, 3),
        "bc_cc_speedup_vs_per_item": round(tbi / tbp, 2),
        "bc_cc_effective_TEPS_persistent": round(n * m / tbp, 1),
        "bc_cc_work_TEPS_persistent": round(work / tbp, 1),
    })
    return res


def extra_reduce(ctx, args):
    """Graph reduction (degree-1 folding + biconnected split + packing), the
    one-time step per graph before reduced closeness / BC: host C
    (cent_reduce, Hopcroft-Tarjan) against GPU kernels (cent_reduce_gpu,
    Tarjan-Vishkin over a BFS forest), input CSR in host memory and result
    back in host memory for both.  Each GPU result is checked against the
    host one (same block / copy / edge / fold counts)."""
    import bc_amd
    out = {}
    graphs = (("blocktree2000", lambda: bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3)),
              ("blocktree30000", lambda: bc_amd.gen_blocktree(30000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=5)),
              ("rmat20", lambda: bc_amd.gen_rmat(20, 16 << 20, seed=2)),
              ("grid2048", lambda: bc_amd.gen_grid(2048, 2048, 0.05, 1000, seed=4)))
    for name, make in graphs:
        g = make()
        host = bc_amd.Reduced(g)
        th = timed(lambda: bc_amd.Reduced(g), 2)
        td = timed(lambda: bc_amd.Reduced(g, ctx=ctx), 3)
        dev = bc_amd.Reduced(g, ctx=ctx)
        for f in ("np", "nnzp", "nblocks", "nfold", "n_ap", "ncomp"):
            assert getattr(dev, f) == getattr(host, f), (name, f)
        out[name] = {"n": g.n, "m": g.m, "blocks": host.nblocks, "articulation": host.n_ap, "folded": host.nfold,
                     "host_ms": round(th * 1e3, 2), "gpu_ms": round(td * 1e3, 2), "speedup": round(th / td, 2)}
        log(f"reduce {name}: n={g.n} m={g.m} blocks={host.nblocks}: host {th * 1e3:.1f} ms, GPU {td * 1e3:.1f} ms")
        del g, host, dev
    return out


def extra_blocktree_bc(ctx, args):
    """All-source Brandes BC + closeness on the block-tree graph (unreduced)."""
    import bc_amd
    g = bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3)
    dg = ctx.upload(g)
    ctx.reset_stats()
    t = timed(lambda: ctx.bc(dg, None), 1)
    st = ctx.stats()
    log(f"blocktree-bc: {t * 1e3:.1f} ms, levels/batch {st['levels'] / max(1, st['batches']):.1f}")
    return {"workload": "all-source BC + closeness, block tree (unreduced)", "n": g.n, "m": g.m,
            "ms": round(t * 1e3, 2), "TEPS": round(g.n * g.m / t, 1),
            "levels_per_batch": round(st["levels"] / max(1, st["batches"]), 2)}


def extra_grid(ctx, args):
    """2048x2048 4-neighbour grid, 5% edges removed, largest component, +1000
    shortcuts, seed 4; 1024 sampled sources, BC + closeness."""
    import bc_amd
    t0 = time.time()
    g = bc_amd.gen_grid(2048, 2048, 0.05, 1000, seed=4)
    src = bc_amd.sample_sources(g.n, 1024, 4)
    log(f"grid: n={g.n} m={g.m} (built in {time.time() - t0:.1f}s)")
    dg = ctx.upload(g)
    ctx.reset_stats()
    t1 = time.perf_counter()
    ctx.bc(dg, src, timing=True)
    t = time.perf_counter() - t1
    st = ctx.stats()
    # bytes model per source (the graph is connected): Brandes 32m + 16 D_s + 60n,
    # plain BFS farness 8m + 12n; the source-parallel step is one launch
    k = len(src)
    b_bc = k * (32 * g.m + 60 * g.n) + 16 * st["dag_edges"]
    kms = st["ms_fwd"]
    log(f"grid: {t * 1e3:.1f} ms, levels/batch {st['levels'] / max(1, st['batches']):.1f}, "
        f"kernel {kms:.1f} ms, {b_bc / (kms * 1e-3) / 1e9:.0f} GB/s by the bytes model")
    ctx.reset_stats()
    t2 = time.perf_counter()
    ctx.cc(dg, src, timing=True)
    tc = time.perf_counter() - t2
    kc = ctx.stats()["ms_fwd"]
    b_cc = k * (8 * g.m + 12 * g.n)
    log(f"grid closeness only: {tc * 1e3:.1f} ms, kernel {kc:.1f} ms, "
        f"{b_cc / (kc * 1e-3) / 1e9:.0f} GB/s by the bytes model")
    return {"workload": "sampled BC + closeness, 2048^2 grid (5% removed, LCC, +1000 shortcuts), "
                        "1024 sources", "n": g.n, "m": g.m, "ms": round(t * 1e3, 2),
            "TEPS": round(k * g.m / t, 1),
            "levels_per_batch": round(st["levels"] / max(1, st["batches"]), 2),
            "kernel": "k_ps_bc", "kernel_ms": round(kms, 2),
            "algorithmic_GBs": round(b_bc / (kms * 1e-3) / 1e9, 1),
            "frac": round(b_bc / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "closeness_only": {"kernel": "k_ps_cc", "ms": round(tc * 1e3, 2), "kernel_ms": round(kc, 2),
                               "TEPS": round(k * g.m / tc, 1),
                               "algorithmic_GBs": round(b_cc / (kc * 1e-3) / 1e9, 1),
                               "frac": round(b_cc / (kc * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def extra_closeness(ctx, args):
    """Closeness alone: the main workload's 4096 sources on R-MAT 2^20 through
    the interleaved step, and all-vertex closeness on the serving pool's first
    R-MAT graph (174k vertices), reduced (D1 + AP, persistent items plus
    bounded two-stream big items) and plain.  Farness must agree."""
    import numpy as np
    import bc_amd

    def best_of(fn, reps):
        fn()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t)
        return min(ts), out

    g = bc_amd.gen_rmat(args.scale, args.edge_factor << args.scale, 0.57, 0.19, 0.19, seed=2)
    src = bc_amd.sample_sources(g.n, args.sources, 2)
    dg = ctx.upload(g)
    t20, (far20, _) = best_of(lambda: ctx.cc(dg, src), 3)
    del dg
    _, gp = serve_pool(5, 12)[0]
    red = bc_amd.Reduced(gp)
    dgp = ctx.upload(gp)
    tr, far_r = best_of(lambda: ctx.cc_reduced(red), 3)
    tp, (far_p, _) = best_of(lambda: ctx.cc(dgp, None), 2)
    log(f"closeness: R-MAT 2^{args.scale} {len(src)} sources {t20 * 1e3:.1f} ms; pool R-MAT (n={gp.n}) all-vertex "
        f"reduced {tr * 1e3:.1f} ms, plain {tp * 1e3:.1f} ms")
    return {"workload": "closeness only: R-MAT 2^20 with the main workload's sources; all-vertex "
                        "closeness on the serving pool's R-MAT graph 0 (seed 5000)",
            "rmat20_sources": len(src), "rmat20_ms": round(t20 * 1e3, 2),
            "rmat20_TEPS": round(len(src) * g.m / t20, 1),
            "pool_rmat_n": gp.n, "pool_rmat_m": gp.m,
            "reduced_ms": round(tr * 1e3, 2), "plain_ms": round(tp * 1e3, 2),
            "reduced_effective_TEPS": round(gp.n * gp.m / tr, 1),
            "reduced_work_TEPS": round(reduced_work_edges(red) / tr, 1),
            "plain_TEPS": round(gp.n * gp.m / tp, 1),
            "farness_equal": bool(np.array_equal(far_r, far_p))}


def extra_rmat14(ctx, args):
    """Exact all-source BC + closeness on R-MAT 2^14 (8 edge draws per vertex,
    LCC, seed 1).  The engine's time beside the
    reference's own computeBC and computeCC, compiled from its sources into
    oracle/_ref (when that build exists), run single-threaded on this host on
    the same graph written in the reference's edge-list format.  Results are
    cross-checked: farness exact, per-predecessor BC within the reference's
    FP32 rounding."""
    import subprocess
    import tempfile
    import numpy as np
    import bc_amd
    g = bc_amd.gen_rmat(14, 8 << 14, seed=1)
    dg = ctx.upload(g)
    n, m = g.n, g.m
    ctx.bc(dg, None)  # warm the workspace
    t0 = time.perf_counter()
    bc, far, _ = ctx.bc(dg, None)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    bcc, _, _ = ctx.bc(dg, None, mode=bc_amd.CENT_BC_REF_COMPAT, farness=False)
    t_gpu_compat = time.perf_counter() - t0
    lay = dg.layout.as_dict()
    lay.pop("ms")
    res = {"workload": "exact all-source BC + closeness, R-MAT 2^14, 8 draws per vertex, LCC, seed 1",
           "n": n, "m": m, "sources": n, "ms": round(t_gpu * 1e3, 2),
           "teps": round(n * m / t_gpu, 1), "ms_ref_formula": round(t_gpu_compat * 1e3, 2), "launch": lay}
    ref_bc = os.path.join(ROOT, "oracle", "_ref", "ref_bc")
    ref_cc = os.path.join(ROOT, "oracle", "_ref", "ref_cc")
    if os.path.exists(ref_bc) and os.path.exists(ref_cc):
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "rmat14.txt")
            bc_amd.write_edgelist(path, g)
            t0 = time.perf_counter()
            subprocess.run([ref_bc, path, os.path.join(td, "bc.txt")], check=True, timeout=900,
                           stdout=subprocess.DEVNULL)
            t_rbc = time.perf_counter() - t0
            t0 = time.perf_counter()
            subprocess.run([ref_cc, "cc", path, os.path.join(td, "cc.txt")], check=True, timeout=900,
                           stdout=subprocess.DEVNULL)
            t_rcc = time.perf_counter() - t0
            rbc = np.loadtxt(os.path.join(td, "bc.txt"))
            rcc = np.loadtxt(os.path.join(td, "cc.txt"), dtype=np.int64)
        scale = max(1.0, float(np.abs(bcc).max()))
        res["cpu_baseline"] = {
            "kind": "reference", "cores": 1, "unit": "TEPS",
            "value": round(n * m / (t_rbc + t_rcc), 1),
            "sample": f"whole workload: reference computeBC {t_rbc:.2f} s + computeCC {t_rcc:.2f} s, "
                      f"CPU {cpu_model()}",
            "speedup_vs_engine": round((t_rbc + t_rcc) / t_gpu, 1)}
        res["parity_vs_reference"] = {
            "farness_equal": bool(np.array_equal(rcc[:, 1], far[rcc[:, 0]])),
            "bc_ref_formula_max_rel": float(np.max(np.abs(rbc - bcc[:len(rbc)])) / scale)}
    log(f"rmat14: {res}")
    return res


def serve_pool(seed: int, size: int):
    """Seeded graph pool of the serving workload: vertex counts log-uniform in
    [1e3, 1e6]; even entries R-MAT with 16 edge draws per vertex (average
    degree 16, the edge factor of the R-MAT workloads; largest component
    kept), odd entries block trees as in the block-tree workload (about 31
    vertices per block with the 30% pendant chains)."""
    import numpy as np
    import bc_amd
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(size):
        n = float(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
        if i % 2 == 0:
            scale = max(10, int(round(np.log2(n))))
            g = bc_amd.gen_rmat(scale, 16 << scale, seed=seed * 1000 + i)
            pool.append(("rmat", g))
        else:
            nb = max(8, int(n / 31))
            g = bc_amd.gen_blocktree(nb, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=seed * 1000 + i)
            pool.append(("blocktree", g))
    return pool


def serve_requests(pool, count: int, seed: int = 5):
    """count requests (graph index, kind, sources): each picks a pool graph
    uniformly; half sampled BC with k log-uniform in [64, 1024], half
    all-vertex closeness."""
    import numpy as np
    import bc_amd
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        j = int(rng.integers(0, len(pool)))
        g = pool[j][1]
        if rng.random() < 0.5:
            k = int(min(g.n, np.exp(rng.uniform(np.log(64), np.log(1024)))))
            reqs.append((j, bc_amd.CENT_Q_BC, bc_amd.sample_sources(g.n, k, 1000 + i)))
        else:
            reqs.append((j, bc_amd.CENT_Q_CC, None))
    return reqs


def run_stream(srv, gids, reqs, rate, drain_s):
    """Open-loop Poisson submission at `rate` requests/s.  A collector thread
    takes the results in submission order as they finish (so finished
    results do not pile up in the server).  Requests still unfinished
    drain_s seconds after the last submission are cancelled.  Returns the
    per-request (kind, latency ms), the wall time from the first submission
    to the last completion, and the number cancelled."""
    import queue
    import threading
    import numpy as np
    rng = np.random.default_rng(55)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    pending = queue.Queue()
    lat, last_end = [], [0.0]
    service = {}  # (graph index, kind) -> [count, summed ms on a slot]
    cancelled = [0]
    deadline = [None]

    def collect():
        while True:
            item = pending.get()
            if item is None:
                return
            rid, kind, j = item
            while True:
                dl = deadline[0]
                left = 1000 if dl is None else int(max(0.0, dl - time.perf_counter()) * 1000)
                info = srv.wait(rid, min(left, 1000))
                if info["state"] in ("done", "failed", "cancelled"):
                    break
                if dl is not None and time.perf_counter() >= dl:
                    srv.cancel(rid)
                    cancelled[0] += 1
                    info = srv.wait(rid)
                    break
            st, _, _, info = srv.result(rid)
            if info["state"] == "done":
                lat.append((kind, info["t_end_ms"] - info["t_submit_ms"]))
                last_end[0] = max(last_end[0], time.perf_counter())
                sv = service.setdefault((j, kind), [0, 0.0])
                sv[0] += 1
                sv[1] += info["t_end_ms"] - info["t_start_ms"]

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i, ((j, kind, src), gap) in enumerate(zip(reqs, gaps)):
        t_next += gap
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        pending.put((srv.submit(gids[j], kind, src), kind, j))
        if (i + 1) % 2000 == 0:
            log(f"serve stream: {i + 1}/{len(reqs)} submitted, {len(lat)} done "
                f"({time.perf_counter() - t0:.1f} s)")
    deadline[0] = time.perf_counter() + drain_s
    pending.put(None)
    while th.is_alive():
        th.join(10.0)
        if th.is_alive():
            log(f"serve stream: draining, {len(lat)} done ({time.perf_counter() - t0:.1f} s)")
    return lat, (last_end[0] or time.perf_counter()) - t0, cancelled[0], service


def _served_pass(srv, gids, reqs):
    """All requests submitted at once (saturation); returns (seconds, results)."""
    t0 = time.perf_counter()
    rids = [srv.submit(gids[j], kind, src) for j, kind, src in reqs]
    got = []
    for rid in rids:
        srv.wait(rid)
        st, bc, far, info = srv.result(rid)
        if st != 0:
            raise RuntimeError(f"served request failed: status {st}")
        got.append((bc, far))
    return time.perf_counter() - t0, got


def _plain_pass(ctx, pool, reqs, resident):
    """The plain form, one request at a time: BC through cent_bc, all-vertex
    closeness through the D1/AP reduction with one level-synchronous launch
    sequence per (component, source block) item (CENT_SCHED_PER_ITEM, the
    launch structure of parallel_CC/para_CC.cu:175-364).  resident=None:
    every request uploads its graph and reduces it on the host; otherwise
    resident[j] = (device graph, reduction) made before the timed pass."""
    import bc_amd
    out = []
    t0 = time.perf_counter()
    for j, kind, src in reqs:
        g = pool[j][1]
        if resident is None:
            dg = ctx.upload(g)
            red = bc_amd.Reduced(g) if kind == bc_amd.CENT_Q_CC else None
        else:
            dg, red = resident[j]
        if kind == bc_amd.CENT_Q_BC:
            bc, far, _ = ctx.bc(dg, src)
            out.append((bc, far))
        else:
            out.append((None, ctx.cc_reduced(red, sched=bc_amd.CENT_SCHED_PER_ITEM)))
        if resident is None:
            del dg, red
    return time.perf_counter() - t0, out


def serving(ctx, args, failures):
    """Serving workload (continuous batching): a seeded pool of graphs, a
    Poisson stream of requests at --serve-rate into a --serve-slots server
    that starts with nothing cached, p50/p99 latency and completed queries/s;
    then, on the stream's first --serve-plain requests at saturation, the
    served form against the plain form, each cold (graphs uploaded and
    reduced on demand) and resident (graphs already on the device), and the
    server's slots against 1.  Every served result is checked against the plain
    form's: farness equal, BC within 1e-9 relative."""
    import numpy as np
    import bc_amd
    t0 = time.time()
    pool = serve_pool(5, args.serve_pool)
    reqs = serve_requests(pool, args.serve_requests)
    log(f"serve: pool n={[g.n for _, g in pool]} built in {time.time() - t0:.1f}s, {len(reqs)} requests")
    budget = 200 << 30
    res = {"workload": f"Poisson {args.serve_rate:g} req/s, {len(reqs)} requests over a pool of {len(pool)} "
                       "graphs (n log-uniform in [1e3, 1e6], half R-MAT avg degree 16, half block tree), "
                       "50% sampled BC k log-uniform in [64, 1024], 50% all-vertex closeness; seed 5",
           "pool_n": [g.n for _, g in pool], "pool_m": [g.m for _, g in pool]}
    srv = bc_amd.Server(slots=args.serve_slots, hbm_budget=budget)
    gids = [srv.add_graph(g) for _, g in pool]
    lat, wall, ncancel, service = run_stream(srv, gids, reqs, args.serve_rate, args.serve_drain)
    st = srv.stats()
    srv.close()
    ms = np.array([x for _, x in lat]) if lat else np.zeros(1)
    stream = {"slots": args.serve_slots, "offered_per_s": args.serve_rate, "completed": len(lat), "of": len(reqs),
              "cancelled": ncancel, "wall_s": round(wall, 3), "queries_per_s": round(len(lat) / wall, 2),
              "p50_ms": round(float(np.percentile(ms, 50)), 2), "p99_ms": round(float(np.percentile(ms, 99)), 2),
              "max_running": st["max_running"], "max_queued": st["max_queued"],
              "graph_uploads": st["graph_uploads"]}
    for kind, name in ((bc_amd.CENT_Q_BC, "bc"), (bc_amd.CENT_Q_CC, "cc")):
        x = np.array([v for kd, v in lat if kd == kind])
        if len(x):
            stream[f"{name}_p50_ms"] = round(float(np.percentile(x, 50)), 2)
            stream[f"{name}_p99_ms"] = round(float(np.percentile(x, 99)), 2)
    # slot time per request by graph and kind: where the capacity goes
    names = {bc_amd.CENT_Q_BC: "bc", bc_amd.CENT_Q_CC: "cc"}
    stream["slot_ms_by_graph"] = {f"{j}:{names[kd]}": [c, round(t / c, 2)]
                                  for (j, kd), (c, t) in sorted(service.items())}
    stream["slot_s_total"] = round(sum(t for _, t in service.values()) / 1e3, 2)
    log(f"serve stream: {stream}")
    res["stream"] = stream

    sub = reqs[:args.serve_plain]
    if not sub:
        return res
    t_pc, ref = _plain_pass(ctx, pool, sub, None)
    log(f"serve plain cold: {len(sub)} requests {t_pc:.2f} s")
    resident = {}
    for j in sorted({j for j, _, _ in sub}):
        g = pool[j][1]
        resident[j] = (ctx.upload(g), bc_amd.Reduced(g))
    _plain_pass(ctx, pool, sub[:2], resident)  # device copies of the reductions
    t_pr, got_pr = _plain_pass(ctx, pool, sub, resident)
    log(f"serve plain resident: {t_pr:.2f} s")
    del resident
    srv = bc_amd.Server(slots=args.serve_slots, hbm_budget=budget)
    gids = [srv.add_graph(g) for _, g in pool]
    t_sc, got_sc = _served_pass(srv, gids, sub)
    t_sw, got_sw = _served_pass(srv, gids, sub)
    srv.close()
    srv = bc_amd.Server(slots=1, hbm_budget=budget)
    gids = [srv.add_graph(g) for _, g in pool]
    _served_pass(srv, gids, sub)
    t_s1, got_s1 = _served_pass(srv, gids, sub)
    srv.close()
    log(f"serve served: cold {t_sc:.2f} s, warm {t_sw:.2f} s, one slot warm {t_s1:.2f} s")
    far_equal, bc_rel = True, 0.0
    for got in (got_pr, got_sc, got_sw, got_s1):
        for (pb, pf), (sb, sf) in zip(ref, got):
            far_equal &= bool(np.array_equal(pf, sf))
            if pb is not None:
                bc_rel = max(bc_rel, float(np.max(np.abs(pb - sb)) / max(1.0, float(np.abs(pb).max()))))
    if not far_equal:
        failures.append("serving: a served farness differs from the plain form's")
    if not bc_rel <= 1e-9:
        failures.append(f"serving: served BC differs from the plain form's (max rel {bc_rel:.2e})")
    q = len(sub)
    res["saturation"] = {
        "requests": q,
        "plain_cold_qps": round(q / t_pc, 2), "plain_resident_qps": round(q / t_pr, 2),
        "served_cold_qps": round(q / t_sc, 2), "served_warm_qps": round(q / t_sw, 2),
        "one_slot_warm_qps": round(q / t_s1, 2),
        "served_cold_vs_plain_cold": round(t_pc / t_sc, 2),
        "served_warm_vs_plain_resident": round(t_pr / t_sw, 2),
        "slots": args.serve_slots, "slots_vs_one": round(t_s1 / t_sw, 2),
        "farness_equal": far_equal, "bc_max_rel_diff": bc_rel}
    log(f"serve saturation: {res['saturation']}")
    return res


EXTRA = {"rmat14": extra_rmat14, "blocktree": extra_blocktree, "blocktree-bc": extra_blocktree_bc, "grid": extra_grid,
         "closeness": extra_closeness, "reduce": extra_reduce}

# Sources whose text decides the dense step's HBM traffic per launch: its
# kernels, the level-loop section of levels.hip that launches them, the
# batch grouping (the sort of a call's sources and the landmark keys the
# upload computes for it), and the build flags.  The counter figures in
# profiles/rmat20_traffic.json are used only while their key (this hash, the
# launch layout, the workload) matches the tree.
TRAFFIC_SOURCES = ("bc_amd/csrc/msbfs_kernels.h", "bc_amd/csrc/wave.h", "bc_amd/Makefile")
TRAFFIC_SECTIONS = (
    ("bc_amd/csrc/levels.hip", "// ------------------------------------------------------------ level loop",
     "// ------------------------------------------------------------ frontier-queue mode"),
    ("bc_amd/csrc/levels.hip", "    // Batch order of the sources:", "    // pinned staging: device-order sources in"),
    ("bc_amd/csrc/engine.hip", "// ------------------------------------------------------------ grouping key",
     'extern "C" int cent_graph_create('),
)
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "rmat20_traffic.json")


def kernel_sha16() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in TRAFFIC_SOURCES:
        h.update(f.encode())
        h.update(open(os.path.join(ROOT, f), "rb").read())
    for f, begin, end in TRAFFIC_SECTIONS:
        text = open(os.path.join(ROOT, f)).read()
        h.update(text[text.index(begin):text.index(end)].encode())
    return h.hexdigest()[:16]


def traffic_key(args, layout) -> dict:
    return {"scale": args.scale, "edge_factor": args.edge_factor, "sources": args.sources,
            "launch": {k: layout[k] for k in ("hub_deg", "chunk_edges", "grid_waves", "frontier", "push_pct")},
            "kernel_sha16": kernel_sha16()}


def _read_counters(d: str, counter: str = None):
    """Per-kernel (dispatches, summed value) of one counter from a rocprofv3
    --pmc output directory (its SQLite *_results.db, or a
    *counter_collection.csv).  counter None: the only counter of the run."""
    import csv
    import glob
    import sqlite3
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True):
        con = sqlite3.connect(f)
        rows += con.execute("select kernel_name, dispatch_id, counter_name, value from counters_collection").fetchall()
        con.close()
    if not rows:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows += [(r["Kernel_Name"], r.get("Dispatch_Id", ""), r.get("Counter_Name", ""), float(r["Counter_Value"]))
                     for r in csv.DictReader(open(f))]
    if not rows:
        raise SystemExit(f"no counter results under {d}")
    out = {}
    for name, disp, cname, val in rows:
        if counter and cname != counter:
            continue
        ds, tot = out.get(name, (set(), 0.0))
        ds.add(disp)
        out[name] = (ds, tot + float(val))
    return {k: (len(v[0]), v[1]) for k, v in out.items()}


def write_traffic(args) -> int:
    """--write-traffic FETCH_DIR WRITE_DIR: turns two rocprofv3 --pmc runs of
    this bench (one for FETCH_SIZE and TCC_EA0_RDREQ_DRAM_32B_sum, one for
    WRITE_SIZE, each over `--steps 1 --warmup 0 --no-cpu-baseline --no-serve
    --timed-only --copy-probe`) into profiles/rmat20_traffic.json.  Readings
    are scaled so that the copy probe's reads and writes equal its known
    bytes.  FETCH_SIZE counts every L2 miss sent to the fabric (Infinity-
    Cache hits included); the DRAM counter counts the read requests that
    the fabric routes to the memory controllers."""
    import bc_amd
    fdir, wdir = args.write_traffic
    fetch, write = _read_counters(fdir, "FETCH_SIZE"), _read_counters(wdir, "WRITE_SIZE")
    try:
        dram = _read_counters(fdir, "TCC_EA0_RDREQ_DRAM_32B_sum")
    except SystemExit:
        dram = {}
    probe = [k for k in fetch if "k_copy_probe" in k]
    if not probe:
        raise SystemExit("the counter runs need --copy-probe (calibration kernel)")
    nprobe, fprobe = fetch[probe[0]]
    _, wprobe = write[probe[0]]
    known_kib = (1 << 30) / 1024.0  # 1 GiB read and 1 GiB written per probe launch
    fs = known_kib / (fprobe / nprobe)
    ws = known_kib / (wprobe / nprobe)
    ds = None  # DRAM counter: 32-byte units, calibrated like the others
    if dram.get(probe[0], (0, 0.0))[1] > 0:
        ds = (1 << 30) / (dram[probe[0]][1] / dram[probe[0]][0])
    passes = {"fwd_pass": ("k_fwd", "k_push1"), "bwd_pass": ("k_bwd",)}
    if not args.no_tune_table and os.path.exists(bc_amd.TUNE_TABLE):
        bc_amd.tune_table_load()
    g = bc_amd.gen_rmat(args.scale, args.edge_factor << args.scale, 0.57, 0.19, 0.19, seed=2)
    ctx = bc_amd.Context(0)
    layout = ctx.upload(g).layout.as_dict()
    out = {"workload": f"sampled BC + closeness, R-MAT 2^{args.scale} avg degree {args.edge_factor} (LCC), "
                       f"{args.sources} sources",
           "key": traffic_key(args, layout), "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
           "dram_scale_bytes_per_unit": round(ds, 4) if ds else None,
           "source": "profiles/rmat20_sampled_bc_hbm_counters.md"}
    for pname, prefixes in passes.items():
        kern = {}
        for k, (nd, v) in fetch.items():
            short = k.split("(")[0].replace("void ", "").replace("cent::", "")
            if any(short.startswith(p) for p in prefixes):
                kern[short] = {"dispatches": nd, "fetch_bytes": int(v * fs * 1024),
                               "write_bytes": int(write.get(k, (0, 0.0))[1] * ws * 1024)}
                if ds:
                    kern[short]["dram_read_bytes"] = int(dram.get(k, (0, 0.0))[1] * ds)
        # one level launch of the pass = one dispatch of its main kernel
        main = "k_bwd<false>" if pname == "bwd_pass" else "k_fwd<false>"
        nl = kern[main]["dispatches"] + (kern.get("k_push1_or", {}).get("dispatches", 0) if pname == "fwd_pass" else 0)
        fb = sum(x["fetch_bytes"] for x in kern.values())
        wb = sum(x["write_bytes"] for x in kern.values())
        out[pname] = {"launches_per_step": nl, "fetch_bytes_per_step": fb, "write_bytes_per_step": wb,
                      "traffic_bytes_per_launch": int((fb + wb) / nl), "kernels": kern}
        if ds:
            db = sum(x.get("dram_read_bytes", 0) for x in kern.values())
            out[pname]["dram_read_bytes_per_step"] = db
            out[pname]["dram_bytes_per_launch"] = int((db + wb) / nl)
    json.dump(out, open(args.traffic_out, "w"), indent=1)
    print(json.dumps(out, indent=1))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="sources of the all-thread CPU baseline (default: 8 per thread)")
    ap.add_argument("--cpu-sample-1t", type=int, default=16,
                    help="sources of the single-thread CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch settings on the first 256 sources before timing")
    ap.add_argument("--tune-save", default="", help="with --tune: write the tuning table here")
    ap.add_argument("--no-tune-table", action="store_true",
                    help="ignore the bundled tuning table (engine defaults)")
    ap.add_argument("--keep-order", action="store_true",
                    help="keep the generator's vertex ids on the device (no degree relabel)")
    ap.add_argument("--one-stream", action="store_true",
                    help="timed steps with every launch group on one stream")
    ap.add_argument("--timed-only", action="store_true",
                    help="skip the host-output and one-stream steps run after the timed ones "
                         "(for profiler runs that should cover exactly the timed steps)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM ceiling, counter calibration)")
    ap.add_argument("--extra", default="",
                    help="comma list of extra workloads reported under 'workloads': "
                         "blocktree (reduced closeness, persistent vs per-item launches), "
                         "blocktree-bc (plain BC+CC on the block tree), closeness (closeness-only "
                         "timings), grid (sampled BC+CC on the 2048^2 grid), reduce (graph "
                         "reduction, host vs GPU), rmat14 (exact all-source BC+CC beside the "
                         "reference's own code)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=12)
    ap.add_argument("--serve-slots", type=int, default=12, help="slots of the stream's server")
    ap.add_argument("--serve-plain", type=int, default=60,
                    help="requests of the stream timed at saturation in the plain and served forms")
    ap.add_argument("--serve-drain", type=float, default=240.0,
                    help="seconds after the last submission; unfinished requests are cancelled")
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the traffic file from two rocprofv3 --pmc runs and exit")
    ap.add_argument("--traffic-out", default=TRAFFIC_JSON,
                    help="where --write-traffic writes (default profiles/rmat20_traffic.json)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this "
              "build (sources shard across GPUs with one all-reduce of the BC vector — "
              "design only).", file=sys.stderr)
        return 2
    if args.write_traffic:
        return write_traffic(args)
    # The request server runs each slot on its own HIP stream.  The runtime
    # spreads a process's streams over GPU_MAX_HW_QUEUES hardware queues (4
    # unless the environment sets more), and streams that share a queue run
    # one after another, so each slot gets a queue of its own.  Read by the
    # runtime at its first call, which comes after this.
    if not args.no_serve and args.serve_slots > int(os.environ.get("GPU_MAX_HW_QUEUES", "4")):
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.serve_slots, 32))

    import numpy as np
    import bc_amd

    failures = []
    t0 = time.time()
    g = bc_amd.gen_rmat(args.scale, args.edge_factor << args.scale, 0.57, 0.19, 0.19, seed=2)
    src = bc_amd.sample_sources(g.n, args.sources, 2)
    log(f"graph: R-MAT scale {args.scale} ef {args.edge_factor}: n={g.n} m={g.m} "
        f"(generated in {time.time() - t0:.1f}s)")

    # launch settings: the bundled MI355X tuning table (applied at upload by
    # graph signature), or a fresh tuning run on the first 256 sources
    if not args.no_tune_table and os.path.exists(bc_amd.TUNE_TABLE):
        bc_amd.tune_table_load()
    ctx = bc_amd.Context(0)
    dg = ctx.upload(g, keep_order=args.keep_order)
    if args.tune:
        t1 = time.perf_counter()
        best = ctx.tune(dg, src[:256], bc_amd.CENT_TUNE_BC)
        log(f"tuned in {time.perf_counter() - t1:.1f}s: {best.as_dict()}")
        if args.tune_save:
            bc_amd.tune_table_save(args.tune_save)
    layout = dg.layout.as_dict()
    layout.pop("ms")
    # the step's outputs stay in HBM (device-output call)
    d_bc = ctx.device_array(g.n, np.float64)
    d_far = ctx.device_array(len(src), np.int64)
    d_ds = ctx.device_array(g.n, np.int64)
    for i in range(args.warmup):
        ctx.bc_device(dg, src, d_bc.zero(), d_far, d_ds.zero(), one_stream=args.one_stream)
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.reset_stats()
    times = []
    for i in range(args.steps):
        d_bc.zero()
        d_ds.zero()
        t1 = time.perf_counter()
        ctx.bc_device(dg, src, d_bc, d_far, d_ds, timing=True, one_stream=args.one_stream)
        times.append(time.perf_counter() - t1)
        log(f"step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    st = ctx.stats()
    ws_bytes = ctx.workspace_bytes()
    far, bc_dev, ds_dev = d_far.numpy(), d_bc.numpy(), d_ds.numpy()
    t_host = t_one = st1 = None
    if not args.timed_only:
        # the same step through the host-array call (scores copied back over PCIe)
        bc = np.zeros(g.n)
        t1 = time.perf_counter()
        _, far_h, ds_h = ctx.bc(dg, src, bc=bc, dist_sum=True, one_stream=args.one_stream)
        t_host = time.perf_counter() - t1
        if not (np.array_equal(far_h, far) and np.array_equal(ds_h, ds_dev) and np.array_equal(bc, bc_dev)):
            failures.append("host-output and device-output results differ")
        log(f"host-output step: {t_host * 1e3:.1f} ms")
        # one more step with every launch on one stream: the level passes' own
        # per-launch durations without a second stream sharing the GPU
        ctx.reset_stats()
        d_bc.zero()
        d_ds.zero()
        t1 = time.perf_counter()
        ctx.bc_device(dg, src, d_bc, d_far, d_ds, timing=True, one_stream=True)
        t_one = time.perf_counter() - t1
        st1 = ctx.stats()
        b1 = d_bc.numpy()
        if not (np.array_equal(d_far.numpy(), far) and np.array_equal(d_ds.numpy(), ds_dev)):
            failures.append("one-stream farness / distance sums differ from the two-stream step's")
        rel1 = float(np.max(np.abs(b1 - bc_dev))) / max(1.0, float(np.abs(bc_dev).max()))
        if not rel1 <= 1e-12:
            failures.append(f"one-stream BC differs from the two-stream step's (max rel {rel1:.2e})")
        log(f"one-stream step: {t_one * 1e3:.1f} ms")
    K = args.steps
    t_step = sum(times) / K
    k, n, m = len(src), g.n, g.m
    teps = k * m / t_step
    dag = st["dag_edges"] / K
    pairs = st["reached"] / K
    b_fwd = k * (16 * m + 20 * n)
    b_bwd = k * (16 * m + 40 * n) + 16 * dag

    # Roofline of the dominant level pass, from the one-stream step (on two
    # streams every launch shares the GPU with the other stream's, so its
    # duration is not the kernel's own).  achieved = HBM-side bytes per
    # launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same pass, calibrated
    # against a copy kernel; profiles/rmat20_traffic.json, used only while
    # its key matches this tree and workload) over the pass's average
    # one-stream launch duration (HIP events on the launching stream).
    roofline = {"bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "streams_timed": 1 if args.one_stream else 2}
    stats_for = st1 if st1 is not None else (st if args.one_stream else None)
    if stats_for is not None:
        fms, bms = stats_for["ms_fwd"], stats_for["ms_bwd"]
        per = K if stats_for is st else 1
        fl, bl = stats_for["launches_fwd"] / per, stats_for["launches_bwd"] / per
        dom = "bwd_pass" if bms >= fms else "fwd_pass"
        ms_launch = (bms / per / bl) if dom == "bwd_pass" else (fms / per / fl)
        launches = bl if dom == "bwd_pass" else fl
        b_model = b_bwd if dom == "bwd_pass" else b_fwd
        roofline.update({
            "kernel": ("k_bwd + k_bwd_hubseg + k_bwd_hubfin (Brandes backward level)" if dom == "bwd_pass"
                       else "k_fwd + k_fwd_hubseg + k_fwd_hubfin (Brandes forward level)"),
            "launch_ms": round(ms_launch, 4), "launches_per_step": launches,
            "one_stream_fwd_ms": round(fms / per, 3), "one_stream_bwd_ms": round(bms / per, 3),
            # the per-source bytes model (every edge charged once per source);
            # a 64-lane launch reads an edge once for 64 sources, so this
            # effective rate exceeds any physical bandwidth
            "model_effective_GBs": round(b_model / launches / (ms_launch * 1e-3) / 1e9, 1)})
        tr = json.load(open(TRAFFIC_JSON)) if os.path.exists(TRAFFIC_JSON) else {}
        key = traffic_key(args, layout)
        if tr.get("key") == key and not args.keep_order:
            tb = tr[dom]["traffic_bytes_per_launch"]
            ach = tb / (ms_launch * 1e-3) / 1e9
            tstep = sum(tr[p]["fetch_bytes_per_step"] + tr[p]["write_bytes_per_step"] for p in ("fwd_pass", "bwd_pass"))
            if tr[dom].get("dram_bytes_per_launch"):
                dgbs = tr[dom]["dram_bytes_per_launch"] / (ms_launch * 1e-3) / 1e9
                roofline.update({"dram_GBs": round(dgbs, 1), "dram_frac": round(dgbs / HBM_PEAK_GBS, 4),
                                 "dram_bytes_per_launch": tr[dom]["dram_bytes_per_launch"]})
            roofline.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tb,
                             "basis": "counter bytes per launch / one-stream launch duration",
                             "traffic_bytes_per_step": int(tstep),
                             "step_traffic_GBs": round(tstep / t_step / 1e9, 1),
                             "step_traffic_frac": round(tstep / t_step / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic_source": tr["source"]})
        else:
            # no counters for this tree: a lower bound of the bytes a backward
            # (forward) level must move — every vertex's frontier record once,
            # and per reached (vertex, source) pair its path count read and
            # its coefficient written and gathered once
            lb = (16 * n + 24 * pairs / launches)
            ach = lb / (ms_launch * 1e-3) / 1e9
            roofline.update({"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                             "basis": "lower-bound bytes model per launch (no counters for this tree)"})
    if args.copy_probe:
        roofline["copy_probe_GBs"] = round(ctx.probe_copy_bandwidth(1 << 30, 10), 1)
    else:
        # ceilings of this access pattern on this GPU, measured now: a plain
        # copy, random gathers of whole 128-byte lines and of 64-byte halves
        # (the fabric moves whole lines, so a 64-byte slice of a sigma row
        # costs a line), and whole lines read 1, 2 or 4 to a load
        # instruction (the slice gathers touch 1.3 lines a load in the
        # forward pass and 2.8 in the backward)
        roofline["ceilings_GBs"] = {"copy": round(ctx.probe_copy_bandwidth(1 << 30, 10), 1),
                                    "gather_128B": round(ctx.probe_gather_bandwidth(128, 1 << 30, 10), 1),
                                    "gather_64B": round(ctx.probe_gather_bandwidth(64, 1 << 30, 10), 1),
                                    "lines_per_load": {str(c): round(ctx.probe_gather_lines(c, 1 << 30, 10), 1)
                                                       for c in (1, 2, 4)}}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        # the box's CPU share per GPU is 16 threads, whatever nproc shows
        threads = max(1, min(16, os.cpu_count() or 1, O.max_threads()))
        ks = args.cpu_sample if args.cpu_sample > 0 else 8 * threads
        sub = src[:ks]
        log(f"cpu baseline: oracle Brandes, {ks} sources, {threads} threads")
        t1 = time.perf_counter()
        obc, ofar, ods, _ = O.brandes(g, sub, threads=threads)
        tc = time.perf_counter() - t1
        k1 = min(args.cpu_sample_1t, ks)
        t1 = time.perf_counter()
        O.brandes(g, src[:k1], threads=1)
        tc1 = time.perf_counter() - t1
        # parity of the sample: the timed step's farness of these sources, and
        # a GPU call on the sample itself for BC and distance sums
        gbc, gfar, gds = ctx.bc(dg, sub, dist_sum=True)
        if not (np.array_equal(ofar, far[:ks]) and np.array_equal(ofar, gfar)):
            failures.append("CPU and GPU farness differ")
        if not np.array_equal(ods, gds):
            failures.append("CPU and GPU distance sums differ")
        brel = float(np.max(np.abs(gbc - obc))) / max(1.0, float(np.abs(obc).max()))
        if not brel <= 1e-9:
            failures.append(f"CPU and GPU BC differ (max rel {brel:.2e})")
        cpu = {"value": round(ks * m / tc, 1), "unit": "TEPS", "cores": threads, "kind": "port",
               "sample": f"first {ks} of the {k} sources (same graph), {tc:.2f} s on {threads} threads; "
                         f"CPU {cpu_model()}, nproc {os.cpu_count()}",
               "single_thread": {"value": round(k1 * m / tc1, 1), "sources": k1, "s": round(tc1, 3)},
               "parity": {"farness_equal": bool(np.array_equal(ofar, gfar)), "bc_max_rel": brel}}

    workloads = {}
    for name in [x for x in args.extra.split(",") if x]:
        if name == "serve":
            continue  # the serving workload runs by default (--no-serve skips it)
        workloads[name] = EXTRA[name](ctx, args)
    serve = None if args.no_serve else serving(ctx, args, failures)

    out = {
        "metric": "TEPS (sources×edges/s) BC+CC, batched/served vs plain; p50/p99 query latency",
        "value": round(teps, 1), "unit": "TEPS", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"sampled BC + closeness, R-MAT 2^{args.scale} avg degree "
                               f"{args.edge_factor} (LCC), {k} sources",
                   "n": n, "m": m, "sources": k, "seed": 2, "global_batch": k,
                   "levels_per_batch": round(st["levels"] / max(1, st["batches"]), 2),
                   "dag_edges_per_step": int(dag), "parallelism": "single GPU",
                   "device_workspace_bytes": ws_bytes,
                   "launch": layout},
        "roofline": roofline, "cpu_baseline": cpu,
        "pcie_inclusive": ({"ms_per_step": round(t_host * 1e3, 3), "value": round(k * m / t_host, 1),
                            "unit": "TEPS", "note": "host-array outputs (cent_bc), one step"}
                           if t_host is not None else None),
        "serving": serve,
        "parity": {"ok": not failures, "failures": failures},
    }
    if workloads:
        out["workloads"] = workloads
    print(json.dumps(out), flush=True)
    if failures:
        for f in failures:
            print(f"bench.py: PARITY FAILURE: {f}", file=sys.stderr, flush=True)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
