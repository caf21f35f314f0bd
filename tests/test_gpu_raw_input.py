"""Raw (multigraph) input on the GPU: repeated edges and self-loops kept as
the reference's loader keeps them (CENT_CSR_RAW).  Every level step must give
the reference's own results on the multigraph fixture (computeCC exact,
computeBC's FP32 per-predecessor scores to their rounding; ref_golden.json)
and the oracle's on larger random multigraphs (farness and distance sums
bit-exact, BC within 1e-9)."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

STEPS = ["dense", "sparse", "source"]


def random_multigraph(n, m, seed, hubs=True):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    extra = [u[: m // 5], v[: m // 5]]  # every fifth edge twice more
    if hubs:  # a few high-degree vertices with parallel edges into them (hub rows, level-1 push)
        h = rng.integers(0, n, 8)
        t = rng.integers(0, n, 4000)
        extra = [np.concatenate([extra[0], np.repeat(h, 500), np.repeat(h, 500)]),
                 np.concatenate([extra[1], t, t])]
    loops = np.arange(0, n, 11)
    uu = np.concatenate([u, extra[0], extra[0], loops])
    vv = np.concatenate([v, extra[1], extra[1], loops])
    return bc_amd.CSR.from_edges(n, uu, vv, raw=True)


@pytest.mark.parametrize("frontier", STEPS)
def test_fixture_matches_reference(ctx, frontier):
    gd = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))["multi.txt"]
    g = bc_amd.buildGraph(O.dataset_path("multi.txt"), raw=True)
    dg = ctx.upload(g)
    src = g.sources()
    bc, far, _ = ctx.bc(dg, src, mode=bc_amd.CENT_BC_REF_COMPAT, frontier=frontier)
    assert far.tolist() == gd["cc"]
    np.testing.assert_allclose(bc, gd["bc_f32"], rtol=2e-6, atol=2e-6)
    obc, _, _, _ = O.brandes(g, src, mode=1)
    np.testing.assert_allclose(bc, obc, rtol=1e-12, atol=1e-12)
    cfar, _ = ctx.cc(dg, src, frontier=frontier)
    assert cfar.tolist() == gd["cc"]


@pytest.mark.parametrize("frontier", STEPS)
@pytest.mark.parametrize("mode", [bc_amd.CENT_BC_TEXTBOOK, bc_amd.CENT_BC_REF_COMPAT])
def test_random_multigraph_matches_oracle(ctx, frontier, mode):
    g = random_multigraph(3000, 12000, 7)
    print(f"[gpu] raw multigraph n={g.n} nnz={g.nnz} {frontier} mode {mode}", flush=True)
    for order in ("auto", "keep"):
        dg = ctx.upload(g, order=order)
        src = bc_amd.sample_sources(g.n, 200, 9)
        bc, far, ds = ctx.bc(dg, src, mode=mode, frontier=frontier, dist_sum=True)
        obc, ofar, ods, _ = O.brandes(g, src, mode=mode)
        assert np.array_equal(far, ofar) and np.array_equal(ds, ods)
        np.testing.assert_allclose(bc, obc, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(obc).max()))
        cfar, cds = ctx.cc(dg, src, frontier=frontier, dist_sum=True)
        assert np.array_equal(cfar, ofar) and np.array_equal(cds, ods)


def test_all_vertex_closeness_and_shim(ctx):
    """All-vertex closeness (interleaved 64-lane groups, level 1 pushed) and
    the reduced form on a multigraph: the BFS farness (the reference's
    D1/AP routine is wrong on self-loops, see the host test)."""
    g = random_multigraph(5000, 30000, 11)
    dg = ctx.upload(g)
    want = O.farness(g)
    assert np.array_equal(ctx.cc(dg, None)[0], want)
    red = bc_amd.Reduced(g)
    assert np.array_equal(ctx.cc_reduced(red), want)
    assert np.array_equal(ctx.cc_reduced(bc_amd.Reduced(g, ctx=ctx)), want)
    with pytest.raises(bc_amd.CentError) as e:
        ctx.bc_reduced(red)
    assert e.value.status == bc_amd.CENT_E_UNSUPPORTED
    fx = bc_amd.buildGraph(O.dataset_path("multi.txt"), raw=True)
    CCs = np.zeros(fx.n, np.int64)
    bc_amd.compute_D1_AP_CC_shareBased_DegreeOrder(fx, CCs)
    assert CCs.tolist() == O.farness(fx).tolist()
