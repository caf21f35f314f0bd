"""GPU parity at the sizes the benchmark times, against the CPU oracle.

- The timed step itself: R-MAT 2^20 (16 x 2^20 edge draws, LCC, seed 2),
  4096 sources sampled with seed 2, launch settings from the bundled MI355X
  tuning table, launch groups on two streams, outputs in device memory
  (cent_bc_device) — exactly what bench.py times.  Farness and per-vertex
  distance sums must be bit-exact and BC within 1e-9 relative of the oracle's
  Brandes over the same 4096 sources (Seq_Brandes/Brandes.c:224-372 in FP64).
- The block-tree workload at full size (2000 blocks, seed 3): all-source BC
  and farness through the D1/AP-reduced form against the oracle's plain
  Brandes over every vertex.
- The high-diameter workload at full size: the 2048 x 2048 grid (5% of edges
  removed, LCC, 1000 shortcuts, seed 4) with its 1024 seed-4 sources, on the
  launch bench.py times (tuning table loaded, default level step): farness
  and distance sums bit-exact and BC within 1e-9 of the oracle over all 1024
  sources, and the closeness-only call's farness equal.
- A reduced graph whose device copy failed half-way is rebuilt by the next
  call instead of being reused (the failure is injected on the host).
"""
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


def _threads():
    return max(1, min(16, os.cpu_count() or 1, O.max_threads()))


def _bc_close(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.max(np.abs(got - want))) / scale
    return err, err <= BC_RTOL


@pytest.mark.timeout(900)
def test_timed_rmat20_step_matches_oracle():
    g = bc_amd.gen_rmat(20, 16 << 20, 0.57, 0.19, 0.19, seed=2)
    src = bc_amd.sample_sources(g.n, 4096, 2)
    print(f"[gpu] timed step: R-MAT 2^20 n={g.n} m={g.m}, 4096 sources, tuned, two streams", flush=True)
    bc_amd.tune_table_load()
    try:
        ctx = bc_amd.Context(0)
        dg = ctx.upload(g)
        lay = dg.layout
        assert (lay.hub_deg, lay.chunk_edges, lay.grid_waves) == (1024, 512, 32768), lay.as_dict()
        d_bc = ctx.device_array(g.n, np.float64).zero()
        d_far = ctx.device_array(len(src), np.int64)
        d_ds = ctx.device_array(g.n, np.int64).zero()
        ctx.bc_device(dg, src, d_bc, d_far, d_ds)
        bc, far, ds = d_bc.numpy(), d_far.numpy(), d_ds.numpy()
        # the peer stream took part: its workspace exists
        assert ctx.workspace_bytes() > 0
    finally:
        bc_amd.tune_table_clear()
    print(f"[gpu] timed step done; oracle Brandes on {_threads()} threads", flush=True)
    obc, ofar, ods, _ = O.brandes(g, src, threads=_threads())
    assert np.array_equal(far, ofar), np.nonzero(far != ofar)[0][:10]
    assert np.array_equal(ds, ods), np.nonzero(ds != ods)[0][:10]
    err, ok = _bc_close(bc, obc)
    print(f"[gpu] timed step: farness and distance sums exact, BC max rel err {err:.2e}", flush=True)
    assert ok, err


@pytest.mark.timeout(600)
def test_blocktree_2000_reduced_bc_matches_oracle(ctx):
    g = bc_amd.gen_blocktree(2000, 1.5, 8, 50000, 6.0, 0.3, 0.5, seed=3)
    print(f"[gpu] block tree 2000: n={g.n} m={g.m}", flush=True)
    red = bc_amd.Reduced(g)
    obc, ofar, _, _ = O.brandes(g, None, threads=_threads())
    for sched in (bc_amd.CENT_SCHED_PERSISTENT, bc_amd.CENT_SCHED_PER_ITEM):
        bc, far = ctx.bc_reduced(red, sched)
        assert np.array_equal(far, ofar), (sched, np.nonzero(far != ofar)[0][:10])
        err, ok = _bc_close(bc, obc)
        print(f"[gpu] block tree 2000 sched {sched}: BC max rel err {err:.2e}", flush=True)
        assert ok, (sched, err)
    # the GPU-built reduction gives the same scores
    redg = bc_amd.Reduced(g, ctx=ctx)
    bc, far = ctx.bc_reduced(redg)
    assert np.array_equal(far, ofar)
    assert _bc_close(bc, obc)[1]


@pytest.mark.timeout(900)
def test_grid2048_1024_sources_matches_oracle():
    g = bc_amd.gen_grid(2048, 2048, 0.05, 1000, seed=4)
    src = bc_amd.sample_sources(g.n, 1024, 4)
    print(f"[gpu] grid 2048^2 n={g.n} m={g.m}, 1024 sources, tuned, default step", flush=True)
    bc_amd.tune_table_load()
    try:
        ctx = bc_amd.Context(0)
        dg = ctx.upload(g)
        bc, far, ds = ctx.bc(dg, src, dist_sum=True)
        cfar, cds = ctx.cc(dg, src, dist_sum=True)
    finally:
        bc_amd.tune_table_clear()
    print(f"[gpu] grid done; oracle Brandes on {_threads()} threads", flush=True)
    obc, ofar, ods, _ = O.brandes(g, src, threads=_threads())
    assert np.array_equal(far, ofar), np.nonzero(far != ofar)[0][:10]
    assert np.array_equal(ds, ods), np.nonzero(ds != ods)[0][:10]
    assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)
    err, ok = _bc_close(bc, obc)
    print(f"[gpu] grid: farness and distance sums exact, BC max rel err {err:.2e}", flush=True)
    assert ok, err


def test_reduced_device_copy_rebuilt_after_failed_upload(ctx):
    g = bc_amd.gen_blocktree(300, 1.5, 8, 2000, 6.0, 0.3, 0.5, seed=7)
    want = O.farness(g)
    for nth in (1, 3, 5):  # device CSR, then the weights, then the work items
        red = bc_amd.Reduced(g)
        bc_amd.debug_fail_upload(nth)
        try:
            with pytest.raises(bc_amd.CentError) as e:
                ctx.cc_reduced(red)
            assert e.value.status == bc_amd.CENT_E_OOM
        finally:
            bc_amd.debug_fail_upload(0)
        got = ctx.cc_reduced(red)
        assert np.array_equal(got, want), nth
        bc, far = ctx.bc_reduced(red)
        assert np.array_equal(far, want), nth
