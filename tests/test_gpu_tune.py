"""Autotuner on the GPU: any launch setting gives the same results.

Bar: farness and distance sums bit-exact; BC within 1e-9 relative of the
oracle (the hub split changes only the FP64 summation order of hub rows).
"""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu

BC_RTOL = 1e-9


@pytest.fixture(autouse=True)
def clean_table():
    bc_amd.tune_table_clear()
    yield
    bc_amd.tune_table_clear()


def close(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    np.testing.assert_allclose(got, want, rtol=BC_RTOL, atol=BC_RTOL * scale)


@pytest.fixture(scope="module")
def rmat():
    g = bc_amd.gen_rmat(14, 8 << 14, seed=1)
    src = bc_amd.sample_sources(g.n, 150, 7)
    obc, ofar, ods, _ = O.brandes(g, src)
    return g, src, obc, ofar, ods


SETTINGS = [
    dict(hub_deg=32, chunk_edges=64, grid_waves=64, frontier="dense"),
    dict(hub_deg=128, chunk_edges=512, grid_waves=4096, frontier="dense"),
    dict(hub_deg=1024, chunk_edges=4096, grid_waves=32768, frontier="dense"),
    dict(hub_deg=256, chunk_edges=1024, grid_waves=0, frontier="sparse"),
]


@pytest.mark.parametrize("st", SETTINGS, ids=lambda d: "-".join(str(v) for v in d.values()))
def test_layout_invariance(ctx, rmat, st):
    g, src, obc, ofar, ods = rmat
    print(f"[gpu] tune layout {st}", flush=True)
    dg = ctx.upload(g)
    dg.set_layout(bc_amd.tune_params(**st))
    lay = dg.layout
    assert (lay.hub_deg, lay.chunk_edges, lay.grid_waves) == (st["hub_deg"], st["chunk_edges"], st["grid_waves"])
    bc, far, ds = ctx.bc(dg, src, dist_sum=True)
    close(bc, obc)
    assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
    cfar, cds = ctx.cc(dg, src, dist_sum=True)
    assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)


def test_tune_picks_applies_and_records(ctx, rmat):
    g, src, obc, ofar, ods = rmat
    print("[gpu] cent_tune rmat14", flush=True)
    dg = ctx.upload(g)
    best = ctx.tune(dg, src[:64], bc_amd.CENT_TUNE_BC, max_trials=6)
    assert best.ms > 0
    lay = dg.layout
    assert (lay.hub_deg, lay.chunk_edges, lay.grid_waves, lay.frontier) == \
        (best.hub_deg, best.chunk_edges, best.grid_waves, best.frontier)
    assert bc_amd.tune_table_get(dg.signature, bc_amd.CENT_TUNE_BC).hub_deg == best.hub_deg
    assert bc_amd.tune_table_get(dg.signature, bc_amd.CENT_TUNE_CC) is None
    bc, far, _ = ctx.bc(dg, src)
    close(bc, obc)
    assert np.array_equal(far, ofar)
    # a fresh upload of the same graph picks the recorded setting up ...
    dg2 = ctx.upload(g)
    assert dg2.layout.hub_deg == best.hub_deg and dg2.layout.frontier == best.frontier
    # ... unless told not to
    assert dg2.layout_for(bc_amd.CENT_TUNE_CC).frontier == bc_amd.FRONTIER["auto"]
    dg3 = ctx.upload(g, tuned=False)
    assert dg3.layout.hub_deg == 256 and dg3.layout.frontier == bc_amd.FRONTIER["auto"]
    # closeness tuned separately: its level step is kept apart from betweenness'
    cbest = ctx.tune(dg, src[:64], bc_amd.CENT_TUNE_CC, max_trials=4)
    assert dg.layout_for(bc_amd.CENT_TUNE_CC).frontier == cbest.frontier
    assert dg.layout_for(bc_amd.CENT_TUNE_BC).frontier == best.frontier
    far, _ = ctx.cc(dg, src)
    assert np.array_equal(far, ofar)


def test_tune_rejects_bad_arguments(ctx, rmat):
    g, src, *_ = rmat
    dg = ctx.upload(g)
    for bad in (dict(sources=src[:0]), dict(sources=np.arange(300)), dict(sources=[g.n]),
                dict(sources=src[:8], what=3)):
        with pytest.raises(bc_amd.CentError):
            ctx.tune(dg, bad["sources"], bad.get("what", bc_amd.CENT_TUNE_BC))
    with pytest.raises(bc_amd.CentError):
        dg.set_layout(bc_amd.tune_params(hub_deg=16))
    assert dg.layout.hub_deg == 256


def test_tune_sample_limits(ctx, rmat):
    """Betweenness tunes on at most 256 sample sources, closeness on at most
    1024 (two launch groups of eight 64-lane batches)."""
    g = rmat[0]
    dg = ctx.upload(g, tuned=False)
    src = bc_amd.sample_sources(g.n, 1025, 3)
    with pytest.raises(bc_amd.CentError):
        ctx.tune(dg, src[:257], bc_amd.CENT_TUNE_BC, max_trials=1)
    with pytest.raises(bc_amd.CentError):
        ctx.tune(dg, src[:1025], bc_amd.CENT_TUNE_CC, max_trials=1)
    best = ctx.tune(dg, src[:1024], bc_amd.CENT_TUNE_CC, max_trials=2)
    assert best.ms > 0
    far, _ = ctx.cc(dg, src[:1024])
    assert np.array_equal(far, O.farness(g, src[:1024]))
