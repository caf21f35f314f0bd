"""Device-output calls (cent_bc_device / cent_cc_device): the results left in
device memory match the host-output calls bit for bit, with the same
accumulate/assign rules and caller vertex order, and agree with the oracle."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("frontier", ["dense", "sparse", "source"])
@pytest.mark.parametrize("keep_order", [False, True])
def test_bc_device_matches_host(ctx, frontier, keep_order):
    print(f"[gpu] bc_device frontier={frontier} keep_order={keep_order}", flush=True)
    g = bc_amd.gen_rmat(13, 16 << 13, seed=21)
    dg = ctx.upload(g, keep_order=keep_order)
    src = bc_amd.sample_sources(g.n, 300, 5)  # > 64: batches regrouped on the dense step
    hb, hf, hd = ctx.bc(dg, src, bc=np.full(g.n, 0.25), dist_sum=True, frontier=frontier)
    db = ctx.device_array(g.n, np.float64, 0.25)
    df = ctx.device_array(len(src), np.int64, -7)
    dd = ctx.device_array(g.n, np.int64, 3)
    ctx.bc_device(dg, src, db, df, dd, frontier=frontier)
    assert np.array_equal(db.numpy(), hb)   # accumulated onto 0.25, like the host call
    assert np.array_equal(df.numpy(), hf)   # assigned
    assert np.array_equal(dd.numpy() - 3, hd)  # accumulated
    obc, ofar, ods, _ = O.brandes(g, src)
    np.testing.assert_allclose(hb - 0.25, obc, rtol=1e-9, atol=1e-9 * max(1.0, obc.max()))
    assert np.array_equal(hf, ofar) and np.array_equal(hd, ods)


@pytest.mark.parametrize("frontier", ["dense", "sparse", "source"])
def test_cc_device_matches_host(ctx, frontier):
    print(f"[gpu] cc_device frontier={frontier}", flush=True)
    g = bc_amd.gen_rmat(13, 16 << 13, seed=22)
    dg = ctx.upload(g)
    src = bc_amd.sample_sources(g.n, 700, 6)  # two multi-batch launches on the dense step
    hf, hd = ctx.cc(dg, src, dist_sum=True, frontier=frontier)
    df = ctx.device_array(len(src), np.int64).zero()
    dd = ctx.device_array(g.n, np.int64).zero()
    ctx.cc_device(dg, src, df, dd, frontier=frontier)
    assert np.array_equal(df.numpy(), hf)
    assert np.array_equal(dd.numpy(), hd)
    _, ofar, ods, _ = O.brandes(g, src)
    assert np.array_equal(hf, ofar) and np.array_equal(hd, ods)


def test_device_all_sources_and_optional_outputs(ctx):
    """k = -1 (every vertex, farness indexed by vertex) on a fixture, and the
    optional outputs left out."""
    g = bc_amd.buildGraph(O.dataset_path("football.txt"))
    dg = ctx.upload(g)
    hb, hf, _ = ctx.bc(dg, None)
    db = ctx.device_array(g.n, np.float64).zero()
    ctx.bc_device(dg, None, db)
    df = ctx.device_array(g.n, np.int64).zero()
    ctx.bc_device(dg, None, db, df)  # accumulates a second time
    assert np.array_equal(db.numpy(), hb + hb)
    assert np.array_equal(df.numpy(), hf)
    cf = ctx.device_array(g.n, np.int64).zero()
    ctx.cc_device(dg, None, cf)
    assert np.array_equal(cf.numpy(), hf)


def test_device_outputs_rejected(ctx):
    g = bc_amd.gen_rmat(10, 8 << 10, seed=3)
    dg = ctx.upload(g)
    src = [0, 5, 9]
    with pytest.raises(ValueError):
        ctx.bc_device(dg, src, ctx.device_array(g.n, np.float32))
    with pytest.raises(ValueError):
        ctx.bc_device(dg, src, ctx.device_array(g.n - 1, np.float64))
    with pytest.raises(ValueError):
        ctx.cc_device(dg, src, np.zeros(len(src), np.int64))  # host array
    s = np.asarray(src, np.int32)
    sp = s.ctypes.data_as(C.POINTER(C.c_int32))
    # through the C-ABI itself: a host array is not device memory ...
    host = np.zeros(g.n)
    st = bc_amd._lib.cent_bc_device(ctx._h, dg._h, sp, len(s), host.ctypes.data, None, None, None)
    assert st == bc_amd.CENT_E_INVAL
    # ... and an allocation too small for n scores is refused before any launch
    small = ctx.device_array(8, np.float64)
    assert g.n * 8 > 4096  # larger than a page-rounded 64-byte allocation
    st = bc_amd._lib.cent_bc_device(ctx._h, dg._h, sp, len(s), small.ptr, None, None, None)
    assert st == bc_amd.CENT_E_INVAL
    # the context still works after the rejected calls
    df = ctx.device_array(len(src), np.int64).zero()
    ctx.cc_device(dg, src, df)
    assert np.array_equal(df.numpy(), ctx.cc(dg, src)[0])


def test_device_edge_cases(ctx):
    """Empty source lists, a graph with isolated vertices (k = -1), and the
    device-array helpers' argument checks."""
    g = bc_amd.CSR.from_edges(6, [0, 1, 3], [1, 2, 4])  # vertex 5 isolated
    dg = ctx.upload(g)
    db = ctx.device_array(g.n, np.float64, 0.5)
    ctx.bc_device(dg, [], db)                      # nothing to do: bc unchanged
    assert np.array_equal(db.numpy(), np.full(g.n, 0.5))
    df = ctx.device_array(g.n, np.int64, -1)
    ctx.bc_device(dg, None, db, df)                # every vertex a source
    hb, hf, _ = ctx.bc(dg, None, bc=np.full(g.n, 0.5))
    assert np.array_equal(db.numpy(), hb) and np.array_equal(df.numpy(), hf)
    assert df.numpy()[5] == 0
    cf = ctx.device_array(g.n, np.int64).zero()
    ctx.cc_device(dg, None, cf)
    assert np.array_equal(cf.numpy(), hf)
    # helpers: bad copy kind, negative sizes, zero-size operations
    a = ctx.device_array(4, np.int64).zero()
    host = np.arange(4, dtype=np.int64)
    assert bc_amd._lib.cent_device_copy(ctx._h, a.ptr, host.ctypes.data, 32, 7) == bc_amd.CENT_E_INVAL
    assert bc_amd._lib.cent_device_fill(ctx._h, a.ptr, 0, -1) == bc_amd.CENT_E_INVAL
    assert bc_amd._lib.cent_device_copy(ctx._h, a.ptr, host.ctypes.data, 0, bc_amd.CENT_COPY_TO_DEVICE) == 0
    a.copy_from(host)
    assert np.array_equal(a.numpy(), host)
