"""Workspace estimates against what a fresh context really allocates: for
every level step, both stream forms, the reference-compat formula, the
reduced form's schedules and a deep graph whose dense level store grows,
cent_ctx_workspace_bytes after the call must not exceed the estimate made
before the graph was uploaded (cent_workspace_estimate), the estimate of the
uploaded graph (cent_graph_workspace_estimate) or the reduced estimate
(cent_reduced_workspace_estimate).  The estimates are also held to within a
factor of the real footprint where the graph's depth is known, so a bound
that no longer follows the engine shows up here."""
import numpy as np
import pytest

import bc_amd
from conftest import probe_depth

pytestmark = pytest.mark.gpu


def check(g, what, k, tight=4.0, **kw):
    ctx = bc_amd.Context(0)
    depth = probe_depth(g)
    before = bc_amd.workspace_estimate(g, what, k, depth, **kw)
    unknown = bc_amd.workspace_estimate(g, what, k, -1, **kw)
    dg = ctx.upload(g)
    after = dg.workspace_estimate(what, k, **kw)
    src = None if k < 0 else bc_amd.sample_sources(g.n, k, 7)
    mode = kw.get("mode", bc_amd.CENT_BC_TEXTBOOK)
    call = dict(frontier=kw.get("frontier", "auto"), one_stream=kw.get("one_stream", False),
                grid_waves=kw.get("grid_waves", 0))
    if what == bc_amd.CENT_TUNE_BC:
        ctx.bc(dg, src, mode=mode, dist_sum=True, **call)
    else:
        ctx.cc(dg, src, dist_sum=True, **call)
    used = ctx.workspace_bytes()
    label = f"n={g.n} what={what} k={k} {kw}: used {used} before {before} after {after} unknown {unknown}"
    assert used <= before and used <= after, label
    if kw.get("frontier", "auto") == "auto":
        assert used <= unknown, label
    assert after <= tight * used + (64 << 20), label
    return used, before, after


def test_plain_calls_on_every_level_step():
    rmat14 = bc_amd.gen_rmat(14, 8 << 14, seed=1)
    check(rmat14, bc_amd.CENT_TUNE_BC, 256)                                  # source-parallel (small BC)
    check(rmat14, bc_amd.CENT_TUNE_BC, 256, frontier="dense")                # two streams
    check(rmat14, bc_amd.CENT_TUNE_BC, 256, frontier="dense", one_stream=True)
    check(rmat14, bc_amd.CENT_TUNE_BC, 128, frontier="dense", mode=bc_amd.CENT_BC_REF_COMPAT)
    check(rmat14, bc_amd.CENT_TUNE_BC, 64, frontier="sparse")
    check(rmat14, bc_amd.CENT_TUNE_CC, 1024)                                 # interleaved closeness
    check(rmat14, bc_amd.CENT_TUNE_CC, -1)
    check(rmat14, bc_amd.CENT_TUNE_CC, 300, frontier="source")
    rmat16 = bc_amd.gen_rmat(16, 16 << 16, seed=2)
    check(rmat16, bc_amd.CENT_TUNE_BC, 256)                                  # dense, tuned table
    check(rmat16, bc_amd.CENT_TUNE_BC, 256, grid_waves=4096)


def test_deep_graphs():
    grid = bc_amd.gen_grid(128, 128, 0.05, 20, seed=4)
    check(grid, bc_amd.CENT_TUNE_BC, 64)                                     # source-parallel
    check(grid, bc_amd.CENT_TUNE_CC, 256)
    # a path walked by the dense step: its level store grows past its first size
    path = bc_amd.CSR.from_edges(3000, np.arange(2999), np.arange(1, 3000))
    check(path, bc_amd.CENT_TUNE_BC, 64, frontier="dense", one_stream=True)


def test_reduced_schedules():
    tree = bc_amd.gen_blocktree(300, 1.5, 8, 3000, 6.0, 0.3, 0.5, seed=3)
    for sched in (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_SHARED, bc_amd.CENT_SCHED_PER_ITEM):
        ctx = bc_amd.Context(0)
        r = bc_amd.Reduced(tree)
        est = r.workspace_estimate(bc_amd.CENT_TUNE_CC, sched=sched)
        ctx.cc_reduced(r, sched=sched)
        assert ctx.workspace_bytes() <= est, (sched, ctx.workspace_bytes(), est)
    for sched in (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_PER_ITEM):
        ctx = bc_amd.Context(0)
        r = bc_amd.Reduced(tree)
        est = r.workspace_estimate(bc_amd.CENT_TUNE_BC, sched=sched)
        ctx.bc_reduced(r, sched=sched)
        used = ctx.workspace_bytes()
        assert used <= est, (sched, used, est)
        ctx = bc_amd.Context(0)
        src = bc_amd.sample_sources(tree.n, 200, 3)
        est = r.workspace_estimate(bc_amd.CENT_TUNE_BC, k=len(src), sched=sched)
        ctx.bc_reduced_sources(r, src, sched=sched)
        assert ctx.workspace_bytes() <= est, (sched, ctx.workspace_bytes(), est)


def test_gpu_reduction_scratch():
    for g in (bc_amd.gen_rmat(16, 16 << 16, seed=3), bc_amd.gen_blocktree(2000, 1.5, 8, 400, 6.0, 0.3, 0.5, seed=3),
              bc_amd.gen_grid(256, 256, 0.05, 100, seed=4)):
        ctx = bc_amd.Context(0)
        bc_amd.Reduced(g, ctx)
        used, est = ctx.workspace_bytes(), bc_amd.reduce_gpu_workspace_estimate(g)
        assert used <= est, (g.n, g.nnz, used, est)
