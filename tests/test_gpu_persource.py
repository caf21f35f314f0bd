"""Source-parallel betweenness (k_ps_bc) against the oracle on the cases its
slot bookkeeping has to get right:

- epoch-tagged levels: a slot tags every level it writes with the epoch of
  its current source and bumps the epoch per source, so nothing is restored
  between sources; after 32,767 sources a slot resets its tags and starts
  again.  One slot running 33,000 sources crosses that reset.
- record forms: slots whose narrow arrays (16-byte records + an f64 BC
  partial, 24 bytes a vertex) pass 512 MB together use 32-byte records that
  hold the partial.  A 102k-vertex grid at 256 slots (629 MB) runs wide, at
  64 slots (157 MB) narrow; both must give the oracle's results.

Bar: farness and distance sums bit-exact, BC within 1e-9 relative
(Seq_Brandes/Brandes.c:224-372 in FP64).
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


def assert_bc_close(got, want, what=""):
    scale = max(1.0, float(np.abs(want).max()) if len(want) else 1.0)
    np.testing.assert_allclose(got, want, rtol=BC_RTOL, atol=BC_RTOL * scale, err_msg=what)


@pytest.mark.parametrize("mode", [bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT])
def test_one_slot_past_the_epoch_reset(ctx, mode):
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    dg = ctx.upload(g)
    src = np.resize(g.sources(), 33000).astype(np.int32)
    print(f"[gpu] karate, {len(src)} sources on one slot, mode {mode}", flush=True)
    bc, far, ds = ctx.bc(dg, src, mode=mode, dist_sum=True, frontier="source", max_slots=1)
    obc, ofar, ods, _ = O.brandes(g, src, mode=mode, threads=_threads())
    assert np.array_equal(far, ofar), np.nonzero(far != ofar)[0][:10]
    assert np.array_equal(ds, ods)
    assert_bc_close(bc, obc, f"one slot, mode {mode}")
    # the context's next call starts from freshly initialised slots
    bc2, far2, _ = ctx.bc(dg, src[:100], mode=mode, frontier="source", max_slots=1)
    obc2, ofar2, _, _ = O.brandes(g, src[:100], mode=mode)
    assert np.array_equal(far2, ofar2)
    assert_bc_close(bc2, obc2, "after the long call")


def test_wide_and_narrow_records_match_oracle(ctx):
    g = bc_amd.gen_grid(320, 320, 0.05, 20, seed=12)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 300, 12)
    print(f"[gpu] grid 320^2 n={g.n}: 256 slots (wide records) and 64 slots (narrow)", flush=True)
    assert 256 * g.n * 24 > (512 << 20) > 64 * g.n * 24
    bw, fw, dw = ctx.bc(dg, src, dist_sum=True, frontier="source", max_slots=256)
    bn, fn, dn = ctx.bc(dg, src, dist_sum=True, frontier="source", max_slots=64)
    obc, ofar, ods, _ = O.brandes(g, src, threads=_threads())
    for bc, far, ds, what in ((bw, fw, dw, "wide"), (bn, fn, dn, "narrow")):
        assert np.array_equal(far, ofar), what
        assert np.array_equal(ds, ods), what
        assert_bc_close(bc, obc, what)
