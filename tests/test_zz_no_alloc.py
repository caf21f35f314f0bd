"""cent_cc_device_noalloc (the allocation-free closeness call) against
cent_cc_device, through the C-ABI: two calls in a row on one GPU, the second
with a smaller input, farness and distance sums bit-exact; an input too
large for the fixed scratch is refused with the sizes in the error text."""
import numpy as np
import pytest

import _oracle as O
import bc_amd

pytestmark = pytest.mark.gpu


def _run_both(ctx, g, src, relabel=True):
    dg = ctx.upload(g, keep_order=not relabel)
    k = g.n if src is None else len(src)
    far_a, ds_a = ctx.device_array(k, np.int64), ctx.device_array(g.n, np.int64).zero()
    ctx.cc_device(dg, src, far_a, ds_a)
    far_b, ds_b = ctx.device_array(k, np.int64), ctx.device_array(g.n, np.int64, fill=7)
    ctx.cc_device_noalloc(dg, src, far_b, ds_b)
    assert np.array_equal(far_a.numpy(), far_b.numpy())
    assert np.array_equal(ds_a.numpy() + 7, ds_b.numpy())  # accumulated, like cent_cc_device
    far_c = ctx.device_array(k, np.int64)
    ctx.cc_device_noalloc(dg, src, far_c)  # no distance sums
    assert np.array_equal(far_a.numpy(), far_c.numpy())
    return far_a.numpy()


def test_noalloc_matches_cc_device_two_calls(ctx):
    g1 = bc_amd.gen_rmat(16, 16 << 16, seed=41)
    src1 = bc_amd.sample_sources(g1.n, 700, 42)
    print(f"[gpu] noalloc call 1: n={g1.n} k=700", flush=True)
    far1 = _run_both(ctx, g1, src1)
    assert np.array_equal(far1[:64], O.farness(g1, src1[:64]))
    g2 = bc_amd.gen_grid(60, 70, 0.05, 5, seed=43)  # smaller, deep: BFS-order relabel
    print(f"[gpu] noalloc call 2: n={g2.n} k=all", flush=True)
    far2 = _run_both(ctx, g2, None)
    assert np.array_equal(far2, O.farness(g2))
    # unrelabelled graph (distance sums written straight to the output) and a repeat of call 1
    _run_both(ctx, g2, bc_amd.sample_sources(g2.n, 33, 44), relabel=False)
    _run_both(ctx, g1, src1[:5])


def test_noalloc_rejects_inputs_beyond_scratch(ctx):
    n = 6_000_000  # a star: 12 bytes a vertex with distance sums on a relabelled graph > 64 MiB
    leaves = np.arange(1, n, dtype=np.int32)
    g = bc_amd.CSR.from_edges(n, np.zeros(n - 1, np.int32), leaves)
    dg = ctx.upload(g)
    src = np.array([0, n - 1], np.int32)
    far, ds = ctx.device_array(2, np.int64), ctx.device_array(n, np.int64).zero()
    with pytest.raises(bc_amd.CentError) as e:
        ctx.cc_device_noalloc(dg, src, far, ds)
    assert e.value.status == bc_amd.CENT_E_INVAL and "scratch" in str(e.value)
    ctx.cc_device_noalloc(dg, src, far)  # without distance sums it fits
    assert far.numpy().tolist() == [n - 1, 1 + 2 * (n - 2)]
