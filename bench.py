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
undirected edges of the (connected) graph.  roofline = the dominant kernel's
algorithmic bytes / its summed HIP-event duration:
    forward  per source: 16m + 20n bytes
    backward per source: 16m + 40n bytes, plus 16 bytes per shortest-path DAG edge
cpu_baseline = the CPU oracle (a C restatement of the same Brandes pass)
timed on a bounded source sample on this host's cores.
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


def extra_blocktree(ctx, args):
    """Block-tree workload: 2000 blocks (Pareto(1.5) sizes in [8, 50000], ER avg
    degree 6 + Hamiltonian cycle), random tree of shared articulation vertices,
    +30% pendant chains (Geometric(0.5) lengths), seed 3; all-source closeness
    through D1 folding + block split.  Compares one persistent launch over the
    work-item queue with level-synchronous launches per work item."""
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
        "bc_cc_persistent_ms": round(tbp * 1e3, 3), "bc_cc_per_item_ms": round(tbi * 1e3, 3),
        "bc_cc_speedup_vs_per_item": round(tbi / tbp, 2),
        "bc_cc_effective_TEPS_persistent": round(n * m / tbp, 1),
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
    the interleaved 8-batch step, and all-vertex closeness on the serving
    pool's 148k-vertex R-MAT graph, reduced (D1 + AP, persistent items plus
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
    log(f"closeness: R-MAT 2^{args.scale} {len(src)} sources {t20 * 1e3:.1f} ms; 148k R-MAT all-vertex "
        f"reduced {tr * 1e3:.1f} ms, plain {tp * 1e3:.1f} ms")
    return {"workload": "closeness only: R-MAT 2^20 with the main workload's sources; all-vertex "
                        "closeness on the serving pool's R-MAT graph 0 (seed 5000)",
            "rmat20_sources": len(src), "rmat20_ms": round(t20 * 1e3, 2),
            "rmat20_TEPS": round(len(src) * g.m / t20, 1),
            "pool_rmat_n": gp.n, "pool_rmat_m": gp.m,
            "reduced_ms": round(tr * 1e3, 2), "plain_ms": round(tp * 1e3, 2),
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
    """Seeded graph pool: vertex counts log-uniform in [1e3, 1e6]; even entries
    R-MAT (avg degree 16, LCC), odd entries block trees (about 31 vertices per
    block with the 30% pendant chains)."""
    import numpy as np
    import bc_amd
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(size):
        n = float(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
        if i % 2 == 0:
            scale = max(10, int(round(np.log2(n))))
            g = bc_amd.gen_rmat(scale, 8 << scale, seed=seed * 1000 + i)
            pool.append(("rmat", g))
        else:
            nb = max(8, int(n / 31))
            g = bc_amd.gen_blocktree(nb, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=seed * 1000 + i)
            pool.append(("blocktree", g))
    return pool


def run_stream(srv, gids, pool, reqs, rate, deadline_s):
    """Open-loop Poisson submission; returns per-request (kind, latency_ms) and
    the wall time from the first submission to the last completion."""
    import numpy as np
    import bc_amd
    rng = np.random.default_rng(55)
    gaps = rng.exponential(1.0 / rate, len(reqs))
    t0 = time.perf_counter()
    t_next = t0
    rids = []
    for (j, kind, src), gap in zip(reqs, gaps):
        t_next += gap
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        rids.append((srv.submit(gids[j], kind, src), kind))
    lat, done_t = [], []
    end = t0 + deadline_s
    for rid, kind in rids:
        left = int(max(0.0, end - time.perf_counter()) * 1000)
        info = srv.wait(rid, left)
        if info["state"] not in ("done", "failed", "cancelled"):
            srv.cancel(rid)
            info = srv.wait(rid)
        st, _, _, info = srv.result(rid)
        if info["state"] == "done":
            lat.append((kind, info["t_end_ms"] - info["t_submit_ms"]))
            done_t.append(info["t_end_ms"])
    return lat, time.perf_counter() - t0


def extra_serve(ctx, args):
    """Serving workload: Poisson arrivals at 200 requests/s; each request picks a
    graph from the pool; half sampled BC (k log-uniform in [64, 1024]), half
    all-vertex closeness (through the D1/AP-reduced path); seed 5.  Reports
    p50/p99 latency and completed queries/s for the continuous-batching server
    (4 slots) and for a single-slot server (one request at a time)."""
    import numpy as np
    import bc_amd
    pool = serve_pool(5, args.serve_pool)
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(args.serve_requests):
        j = int(rng.integers(0, len(pool)))
        g = pool[j][1]
        if rng.random() < 0.5:
            k = int(min(g.n, np.exp(rng.uniform(np.log(64), np.log(1024)))))
            reqs.append((j, bc_amd.CENT_Q_BC, bc_amd.sample_sources(g.n, k, 1000 + i)))
        else:
            reqs.append((j, bc_amd.CENT_Q_CC, None))
    log(f"serve: pool {[(kind, g.n) for kind, g in pool]}, {len(reqs)} requests")
    res = {"workload": f"Poisson {args.serve_rate:g} req/s, {len(reqs)} requests over a pool of "
                       f"{len(pool)} graphs (n log-uniform in [1e3, 1e6], half R-MAT deg 16, half "
                       "block tree), 50% sampled BC k in [64,1024], 50% all-vertex closeness",
           "pool_n": [g.n for _, g in pool]}
    for label, slots in (("served", 4), ("one_at_a_time", 1)):
        srv = bc_amd.Server(slots=slots, hbm_budget=200 << 30)
        gids = [srv.add_graph(g) for _, g in pool]
        # warm the graph cache and reductions
        for gid in gids:
            srv.wait(srv.submit(gid, bc_amd.CENT_Q_BC, [0]))
        lat, wall = run_stream(srv, gids, pool, reqs, args.serve_rate, args.serve_deadline)
        st = srv.stats()
        srv.close()
        ms = np.array([x for _, x in lat]) if lat else np.zeros(1)
        out = {"completed": len(lat), "of": len(reqs), "wall_s": round(wall, 3),
               "queries_per_s": round(len(lat) / wall, 2),
               "p50_ms": round(float(np.percentile(ms, 50)), 2),
               "p99_ms": round(float(np.percentile(ms, 99)), 2),
               "max_running": st["max_running"], "max_queued": st["max_queued"]}
        for kind, name in ((bc_amd.CENT_Q_BC, "bc"), (bc_amd.CENT_Q_CC, "cc")):
            x = np.array([v for kd, v in lat if kd == kind])
            if len(x):
                out[f"{name}_p50_ms"] = round(float(np.percentile(x, 50)), 2)
                out[f"{name}_p99_ms"] = round(float(np.percentile(x, 99)), 2)
        log(f"serve[{label}]: {out}")
        res[label] = out
    res["speedup_queries_per_s"] = round(res["served"]["queries_per_s"] /
                                         max(1e-9, res["one_at_a_time"]["queries_per_s"]), 2)
    if args.serve_plain > 0:
        res["vs_plain"] = serve_vs_plain(ctx, pool, reqs[:args.serve_plain])
    return res


def serve_vs_plain(ctx, pool, reqs):
    """Saturation throughput on the same requests: the served form (4-slot
    server, device-resident graph and reduction cache, persistent item kernel,
    all requests submitted at once) against the plain per-component-launch
    form (one request at a time; each uploads its graph, BC through cent_bc,
    all-vertex closeness through a fresh host D1/AP reduction and one launch
    per (component, source-block) item).  Results must agree: farness exact,
    BC within 1e-9 relative."""
    import numpy as np
    import bc_amd
    out_plain = []
    t0 = time.perf_counter()
    for j, kind, src in reqs:
        g = pool[j][1]
        dg = ctx.upload(g)
        if kind == bc_amd.CENT_Q_BC:
            bc, far, _ = ctx.bc(dg, src)
            out_plain.append((bc, far))
        else:
            red = bc_amd.Reduced(g)
            out_plain.append((None, ctx.cc_reduced(red, sched=bc_amd.CENT_SCHED_PER_ITEM)))
        del dg
    t_plain = time.perf_counter() - t0
    log(f"serve[plain]: {len(reqs)} requests in {t_plain:.2f} s")
    srv = bc_amd.Server(slots=4, hbm_budget=200 << 30)
    gids = [srv.add_graph(g) for _, g in pool]
    for gid in gids:
        srv.wait(srv.submit(gid, bc_amd.CENT_Q_BC, [0]))
    t0 = time.perf_counter()
    rids = [srv.submit(gids[j], kind, src) for j, kind, src in reqs]
    got = []
    for rid in rids:
        srv.wait(rid)
        st, bc, far, info = srv.result(rid)
        if st != 0:
            raise RuntimeError(f"served request failed: status {st}")
        got.append((bc, far))
    t_srv = time.perf_counter() - t0
    srv.close()
    far_equal, bc_rel = True, 0.0
    for (pb, pf), (sb, sf) in zip(out_plain, got):
        far_equal &= bool(np.array_equal(pf, sf))
        if pb is not None:
            bc_rel = max(bc_rel, float(np.max(np.abs(pb - sb)) / max(1.0, float(np.abs(pb).max()))))
    res = {"requests": len(reqs), "plain_s": round(t_plain, 3), "served_s": round(t_srv, 3),
           "plain_queries_per_s": round(len(reqs) / t_plain, 2),
           "served_queries_per_s": round(len(reqs) / t_srv, 2),
           "speedup": round(t_plain / t_srv, 2), "farness_equal": far_equal,
           "bc_max_rel_diff": bc_rel}
    log(f"serve[vs plain]: {res}")
    return res


EXTRA = {"rmat14": extra_rmat14, "blocktree": extra_blocktree, "blocktree-bc": extra_blocktree_bc, "grid": extra_grid,
         "closeness": extra_closeness, "serve": extra_serve, "reduce": extra_reduce}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="sources for the CPU baseline (default: 8 per core)")
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
                         "timings), grid (sampled BC+CC on "
                         "the 2048^2 grid), serve (request server), reduce (graph reduction, "
                         "host vs GPU)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=12)
    ap.add_argument("--serve-plain", type=int, default=60,
                    help="requests of the stream timed in the plain form against the server")
    ap.add_argument("--serve-deadline", type=float, default=90.0,
                    help="seconds after the first submission; unfinished requests are cancelled")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this "
              "build (sources shard across GPUs with one all-reduce of the BC vector — "
              "design only).", file=sys.stderr)
        return 2

    import numpy as np
    import bc_amd

    t0 = time.time()
    g = bc_amd.gen_rmat(args.scale, args.edge_factor << args.scale, 0.57, 0.19, 0.19, seed=2)
    src = bc_amd.sample_sources(g.n, args.sources, 2)
    print(f"# graph: R-MAT scale {args.scale} ef {args.edge_factor}: n={g.n} m={g.m} "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    # launch settings: the bundled MI355X tuning table (applied at upload by
    # graph signature), or a fresh tuning run on the first 256 sources
    if not args.no_tune_table and os.path.exists(bc_amd.TUNE_TABLE):
        bc_amd.tune_table_load()
    ctx = bc_amd.Context(0)
    dg = ctx.upload(g, keep_order=args.keep_order)
    if args.tune:
        t1 = time.perf_counter()
        best = ctx.tune(dg, src[:256], bc_amd.CENT_TUNE_BC)
        print(f"# tuned in {time.perf_counter() - t1:.1f}s: {best.as_dict()}", file=sys.stderr, flush=True)
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
        print(f"# warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.reset_stats()
    times = []
    for i in range(args.steps):
        d_bc.zero()
        d_ds.zero()
        t1 = time.perf_counter()
        ctx.bc_device(dg, src, d_bc, d_far, d_ds, timing=True, one_stream=args.one_stream)
        times.append(time.perf_counter() - t1)
        print(f"# step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    st = ctx.stats()
    ws_bytes = ctx.workspace_bytes()
    far = d_far.numpy()
    t_host = t_one = st1 = None
    if not args.timed_only:
        # the same step through the host-array call (scores copied back over PCIe)
        bc = np.zeros(g.n)
        t1 = time.perf_counter()
        _, far_h, _ = ctx.bc(dg, src, bc=bc, dist_sum=True, one_stream=args.one_stream)
        t_host = time.perf_counter() - t1
        if not (np.array_equal(far_h, far) and np.array_equal(bc, d_bc.numpy())):
            print("# WARNING: host-output and device-output results differ", file=sys.stderr)
        print(f"# host-output step: {t_host * 1e3:.1f} ms", file=sys.stderr, flush=True)
        # one more step with every launch on one stream: the level passes' own
        # rates without the second stream sharing the GPU (not part of value)
        ctx.reset_stats()
        d_bc.zero()
        d_ds.zero()
        t1 = time.perf_counter()
        ctx.bc_device(dg, src, d_bc, d_far, d_ds, timing=True, one_stream=True)
        t_one = time.perf_counter() - t1
        st1 = ctx.stats()
        print(f"# one-stream step: {t_one * 1e3:.1f} ms", file=sys.stderr, flush=True)
    K = args.steps
    t_step = sum(times) / K
    k, n, m = len(src), g.n, g.m
    teps = k * m / t_step
    dag = st["dag_edges"] / K
    b_fwd = k * (16 * m + 20 * n)
    b_bwd = k * (16 * m + 40 * n) + 16 * dag
    ms_fwd = st["ms_fwd"] / K
    ms_bwd = st["ms_bwd"] / K
    if ms_bwd >= ms_fwd:
        kname, b, ms, launches = "k_bwd (Brandes backward level)", b_bwd, ms_bwd, st["launches_bwd"] / K
    else:
        kname, b, ms, launches = "k_fwd (Brandes forward level)", b_fwd, ms_fwd, st["launches_fwd"] / K
    achieved = b / (ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
        "launches_per_step": launches, "kernel_ms_per_step": round(ms, 3),
        "algorithmic_bytes_per_step": int(b),
        "fwd_ms": round(ms_fwd, 3), "bwd_ms": round(ms_bwd, 3),
        "fwd_GBs": round(b_fwd / (ms_fwd * 1e-3) / 1e9, 1) if ms_fwd else None,
        "bwd_GBs": round(b_bwd / (ms_bwd * 1e-3) / 1e9, 1) if ms_bwd else None,
        # launch groups alternate over two streams, so a pass's launches
        # share the GPU with the other stream's: the per-launch durations
        # above overlap, and the whole step's rate is the aggregate
        "streams": 1 if args.one_stream else 2,
        "aggregate_GBs": round((b_fwd + b_bwd) / t_step / 1e9, 1),
        "aggregate_frac": round((b_fwd + b_bwd) / t_step / 1e9 / HBM_PEAK_GBS, 4),
    }
    if st1 is not None:
        f1, w1 = st1["ms_fwd"], st1["ms_bwd"]
        b1 = b_bwd if w1 >= f1 else b_fwd
        t1k = max(f1, w1)
        roofline["one_stream"] = {
            "ms_per_step": round(t_one * 1e3, 3), "fwd_ms": round(f1, 3), "bwd_ms": round(w1, 3),
            "kernel": "k_bwd (Brandes backward level)" if w1 >= f1 else "k_fwd (Brandes forward level)",
            "achieved": round(b1 / (t1k * 1e-3) / 1e9, 1) if t1k else None,
            "frac": round(b1 / (t1k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if t1k else None,
            "launches_per_step": (st1["launches_bwd"] if w1 >= f1 else st1["launches_fwd"]),
        }

    if args.copy_probe:
        roofline["copy_probe_GBs"] = round(ctx.probe_copy_bandwidth(1 << 30, 10), 1)
    # HBM-side bytes per launch come from the rocprofv3 counter passes of this
    # same workload (profiles/rmat20_traffic.json, FETCH_SIZE / WRITE_SIZE
    # scaled by the copy-kernel calibration); they are not measured live.
    tpath = os.path.join(ROOT, "profiles", "rmat20_traffic.json")
    if (args.scale, args.edge_factor, args.sources) == (20, 16, 4096) and os.path.exists(tpath):
        tr = json.load(open(tpath))
        key = "bwd_pass" if kname.startswith("k_bwd") else "fwd_pass"
        roofline["traffic"] = tr[key]["traffic_bytes_per_launch"]
        roofline["traffic_unit"] = "bytes per launch of the level pass"
        roofline["algorithmic_bytes_per_launch"] = int(b / launches)
        # counter-measured bytes of both passes over the live step time: the
        # fabric-side rate of the whole step (the passes of the two streams
        # overlap, so a per-pass rate would count shared time twice)
        tstep = sum(tr[p]["fetch_bytes_per_step"] + tr[p]["write_bytes_per_step"] for p in ("fwd_pass", "bwd_pass"))
        tgbs = tstep / t_step / 1e9
        roofline["traffic_bytes_per_step"] = int(tstep)
        roofline["traffic_GBs"] = round(tgbs, 1)
        roofline["traffic_frac"] = round(tgbs / HBM_PEAK_GBS, 4)
        roofline["traffic_source"] = tr["source"]

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle as O
        threads = max(1, min(16, os.cpu_count() or 1, O.max_threads()))
        ks = args.cpu_sample if args.cpu_sample > 0 else 8 * threads
        sub = src[:ks]
        print(f"# cpu baseline: oracle Brandes, {ks} sources, {threads} threads", file=sys.stderr,
              flush=True)
        t1 = time.perf_counter()
        obc, ofar, ods, _ = O.brandes(g, sub, threads=threads)
        tc = time.perf_counter() - t1
        if not np.array_equal(ofar, far[:ks]):
            print("# WARNING: CPU and GPU farness differ", file=sys.stderr)
        cpu = {"value": round(ks * m / tc, 1), "unit": "TEPS", "cores": threads, "kind": "port",
               "sample": f"first {ks} of the {k} sources (same graph), {tc:.2f} s, "
                         f"CPU {cpu_model()}"}

    workloads = {}
    for name in [x for x in args.extra.split(",") if x]:
        workloads[name] = EXTRA[name](ctx, args)

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
    }
    if workloads:
        out["workloads"] = workloads
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
