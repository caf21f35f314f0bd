"""Builds the MI355X tuning table for the bundled workloads.

Runs cent_tune (closeness, then betweenness) on the plain form of each workload graph (R-MAT 2^14 and
2^20, the 2048x2048 grid with shortcuts, and R-MAT / block-tree graphs
across the serving pool's size range, up to about 900k vertices), then
writes the process table:

    python -m bc_amd.tune_workloads --out bc_amd/tune_mi355x.txt

The graphs are seeded and synthetic, like the benchmark's.  Each line of the
output keys one graph signature; graphs with nearby signatures share it at
upload (cent_graph_create).
"""
import argparse
import sys
import time

import bc_amd


def workloads(quick: bool):
    yield "rmat14", lambda: bc_amd.gen_rmat(14, 8 << 14, seed=1)
    yield "rmat20", lambda: bc_amd.gen_rmat(20, 16 << 20, seed=2)
    if not quick:
        yield "grid2048", lambda: bc_amd.gen_grid(2048, 2048, 0.05, 1000, seed=4)
    for scale in (10, 12, 14, 16, 18):
        yield f"rmat{scale}_ef16", lambda s=scale: bc_amd.gen_rmat(s, 16 << s, seed=100 + s)
    for nb in (40, 400, 4000, 8000) + (() if quick else (30000,)):
        yield f"blocktree{nb}", lambda b=nb: bc_amd.gen_blocktree(n_blocks=b, seed=3)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--sources", type=int, default=256, help="betweenness sample size (at most 256)")
    ap.add_argument("--cc-sources", type=int, default=1024, help="closeness sample size (at most 1024)")
    ap.add_argument("--quick", action="store_true", help="skip the 4M-vertex grid and the largest block tree")
    ap.add_argument("--only", default="", help="comma-separated workload names to tune (default: all)")
    args = ap.parse_args()
    ctx = bc_amd.Context(0)
    bc_amd.tune_table_clear()
    only = set(x for x in args.only.split(",") if x)
    for name, make in workloads(args.quick):
        if only and name not in only:
            continue
        g = make()
        dg = ctx.upload(g, tuned=False)
        # closeness first: the graph keeps the betweenness chunk/hub layout
        for what, label in ((bc_amd.CENT_TUNE_CC, "cc"), (bc_amd.CENT_TUNE_BC, "bc")):
            ks = args.cc_sources if what == bc_amd.CENT_TUNE_CC else args.sources
            src = bc_amd.sample_sources(g.n, min(ks, g.n), 11)
            t0 = time.perf_counter()
            best = ctx.tune(dg, src, what)
            print(f"{name} {label}: n={g.n} nnz={g.nnz} sig={dg.signature} best={best.as_dict()} "
                  f"({time.perf_counter() - t0:.1f} s)", flush=True)
        del dg
    bc_amd.tune_table_save(args.out)
    print(f"wrote {bc_amd.tune_table_size()} entries to {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
