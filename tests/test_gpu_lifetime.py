"""Device copies outlive the context that made them: a reduced graph's device
cache is built by whichever context calls first and a device graph by the
context that uploaded it, and either may be destroyed before the copy is
used again from another context on the same GPU (the copies keep only the
device id)."""
import gc
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


def test_reduced_device_cache_after_its_context_is_gone():
    g = bc_amd.gen_blocktree(60, 1.5, 8, 400, 6.0, 0.3, 0.5, seed=5)
    r = bc_amd.Reduced(g)                   # host build; the device copy comes with the first call
    a = bc_amd.Context(0)
    far_a = a.cc_reduced(r)
    bc_a, _ = a.bc_reduced(r)
    del a
    gc.collect()
    b = bc_amd.Context(0)
    far_b = b.cc_reduced(r)
    bc_b, _ = b.bc_reduced(r)
    src = bc_amd.sample_sources(g.n, 40, 2)
    bc_s, far_s = b.bc_reduced_sources(r, src)
    assert np.array_equal(far_a, far_b) and np.array_equal(far_b, O.farness(g))
    np.testing.assert_allclose(bc_b, bc_a, rtol=1e-12, atol=1e-12)
    want, wfar, _, _ = O.brandes(g, src)
    np.testing.assert_allclose(bc_s, want, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(want).max()))
    assert np.array_equal(far_s, wfar)


def test_device_graph_after_its_context_is_gone():
    g = bc_amd.gen_rmat(10, 8 << 10, seed=3)
    lib = bc_amd._lib
    ha, hb, dg = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.cent_ctx_create(0, C.byref(ha)) == 0
    assert lib.cent_graph_create(ha, bc_amd._ptr(g.row_ptr, C.c_int32), bc_amd._ptr(g.col, C.c_int32), g.n, g.nnz,
                                 0, C.byref(dg)) == 0
    lib.cent_ctx_destroy(ha)
    assert lib.cent_ctx_create(0, C.byref(hb)) == 0
    try:
        far = np.zeros(g.n, np.int64)
        assert lib.cent_cc(hb, dg, None, -1, bc_amd._ptr(far, C.c_int64), None, None) == 0
        assert np.array_equal(far, O.farness(g))
    finally:
        lib.cent_graph_free(dg)
        lib.cent_ctx_destroy(hb)
