#!/usr/bin/env python3
"""Benchmark: sampled Brandes BC + closeness on R-MAT 2^20 (avg degree 16).

Workload (one step): 4096 uniformly sampled sources (seed 2) on an R-MAT graph
with 2^20 ids and 16 x 2^20 edge draws (a=0.57, b=c=0.19, d=0.05, seed 2),
self-loops and duplicates removed, symmetrised, largest component kept, int32
CSR.  One step computes, for every source, its Brandes dependencies
(accumulated into FP64 BC) and its farness (int64) plus the per-vertex int64
distance sums, all in one fused multi-source pass on the GPU.  The graph and
the source list are resident in HBM before the timed region.

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
    ap.add_argument("--keep-order", action="store_true",
                    help="keep the generator's vertex ids on the device (no degree relabel)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM ceiling, counter calibration)")
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

    ctx = bc_amd.Context(0)
    dg = ctx.upload(g, keep_order=args.keep_order)
    bc = np.zeros(g.n)
    for i in range(args.warmup):
        ctx.bc(dg, src, bc=bc, dist_sum=True)
        print(f"# warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.reset_stats()
    times = []
    for i in range(args.steps):
        bc[:] = 0
        t1 = time.perf_counter()
        _, far, ds = ctx.bc(dg, src, bc=bc, dist_sum=True, timing=True)
        times.append(time.perf_counter() - t1)
        print(f"# step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    st = ctx.stats()
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
    }

    if args.copy_probe:
        roofline["copy_probe_GBs"] = round(ctx.probe_copy_bandwidth(1 << 30, 10), 1)

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

    out = {
        "metric": "TEPS (sources×edges/s) BC+CC, batched/served vs plain; p50/p99 query latency",
        "value": round(teps, 1), "unit": "TEPS", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"sampled BC + closeness, R-MAT 2^{args.scale} avg degree "
                               f"{args.edge_factor} (LCC), {k} sources",
                   "n": n, "m": m, "sources": k, "seed": 2, "global_batch": k,
                   "levels_per_batch": round(st["levels"] / max(1, st["batches"]), 2),
                   "dag_edges_per_step": int(dag), "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
