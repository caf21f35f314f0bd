"""Raw input: graphs loaded as the reference loads them, with repeated edges
and self-loops kept (CENT_CSR_RAW; Lib/AdjList/AdjList.c:76-92 does no
cleaning).  Host-side checks: the loader's rows, the oracle against the
reference's own outputs on the multigraph fixture (tests/golden/ref_golden.json,
written by make_golden.py from the reference compiled from its sources), and
the reduction on a multigraph."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN

RAW = "multi.txt"


def golden():
    return json.load(open(os.path.join(GOLDEN, "ref_golden.json")))[RAW]


def test_raw_loader_keeps_parallel_edges_and_self_loops():
    g = bc_amd.buildGraph(O.dataset_path(RAW), raw=True)
    lines = open(O.dataset_path(RAW)).read().split("\n")
    n, m = map(int, lines[0].split())
    assert g.n == n and g.nnz == 2 * m  # every line gives both directions, a self-loop twice in its row
    row = lambda v: list(g.col[g.row_ptr[v]:g.row_ptr[v + 1]])
    assert row(0) == [1, 1, 3]          # 0-1 twice, 3-0 once
    assert row(2).count(2) == 2         # self-loop 2-2
    assert row(12) == [10, 13, 13]
    h = bc_amd.buildGraph(O.dataset_path(RAW))  # the default cleans both
    assert h.nnz < g.nnz and all(h.col[h.row_ptr[v]:h.row_ptr[v + 1]].tolist().count(v) == 0 for v in range(h.n))


def test_constructors_record_the_graph_kind():
    raw = bc_amd.buildGraph(O.dataset_path(RAW), raw=True)
    assert raw.kind == bc_amd.CENT_CSR_KIND_MULTI
    assert bc_amd.buildGraph(O.dataset_path(RAW)).kind == bc_amd.CENT_CSR_KIND_SIMPLE
    # a self-loop alone, and a repeated edge alone, each make a multigraph
    assert bc_amd.CSR.from_edges(3, [0, 1], [1, 1], raw=True).kind == bc_amd.CENT_CSR_KIND_MULTI
    assert bc_amd.CSR.from_edges(3, [0, 1, 1], [1, 2, 0], raw=True).kind == bc_amd.CENT_CSR_KIND_MULTI
    assert bc_amd.CSR.from_edges(3, [0, 1], [1, 2], raw=True).kind == bc_amd.CENT_CSR_KIND_SIMPLE
    # rows handed over as arrays are scanned once
    arr = bc_amd.CSR.from_arrays(raw.row_ptr, raw.col)
    assert arr.kind == bc_amd.CENT_CSR_KIND_MULTI
    assert bc_amd.CSR.from_arrays(bc_amd.gen_grid(4, 4, 0.0, 0, seed=1).row_ptr,
                                  bc_amd.gen_grid(4, 4, 0.0, 0, seed=1).col).kind == bc_amd.CENT_CSR_KIND_SIMPLE


def test_raw_flag_rejects_lcc():
    with pytest.raises(bc_amd.CentError):
        bc_amd.CSR.from_edges(3, [0, 1], [1, 2], lcc=True, raw=True)


def test_oracle_equals_reference_on_multigraph():
    gd = golden()
    g = bc_amd.buildGraph(O.dataset_path(RAW), raw=True)
    src = g.sources()
    assert np.array_equal(O.farness(g, src), gd["cc"])
    # the reference's FP32 per-predecessor BC, digit for digit, and the FP64 form near it
    f32 = O.brandes_f32(g, src).astype(np.float64)
    np.testing.assert_allclose(f32, gd["bc_f32"], rtol=1e-6, atol=1e-6)
    bc, far, _, _ = O.brandes(g, src, mode=1)
    np.testing.assert_allclose(bc, gd["bc_f32"], rtol=2e-6, atol=2e-6)
    # parallel edges change path counts: the simple graph gives other scores
    h = bc_amd.buildGraph(O.dataset_path(RAW))
    assert not np.allclose(O.brandes(h, src, mode=1)[0], gd["bc_f32"], rtol=1e-3)
    assert np.array_equal(O.farness(h, src), gd["cc"])


def test_reference_d1ap_quirk_on_self_loops():
    """The reference's compute_D1_AP_CC_shareBased_DegreeOrder returns 1 for
    vertices 10 and 11 of the fixture (11 has a self-loop and one neighbour);
    its own computeCC gives 44 and 56, the BFS farness.  The engine's reduced
    form returns the BFS farness."""
    gd = golden()
    assert gd["d1ap"][10] == 1 and gd["d1ap"][11] == 1
    assert [c for i, c in enumerate(gd["cc"]) if i not in (10, 11)] == \
           [c for i, c in enumerate(gd["d1ap"]) if i not in (10, 11)]
    g = bc_amd.buildGraph(O.dataset_path(RAW), raw=True)
    r = bc_amd.Reduced(g)
    D = O.weighted_farness(r.np, r.prp, r.pcol, r.pw)
    assert np.array_equal(r.farness_from(D), gd["cc"])
    with pytest.raises(bc_amd.CentError) as e:
        r.bc_from(np.zeros(r.np))
    assert e.value.status == bc_amd.CENT_E_UNSUPPORTED


def test_random_multigraph_reduction_keeps_farness():
    rng = np.random.default_rng(3)
    n = 400
    u = rng.integers(0, n, 1500)
    v = rng.integers(0, n, 1500)
    u = np.concatenate([u, u[:300], np.arange(0, n, 7)])  # repeated edges and self-loops
    v = np.concatenate([v, v[:300], np.arange(0, n, 7)])
    g = bc_amd.CSR.from_edges(n, u, v, raw=True)
    assert g.nnz == 2 * len(u)
    r = bc_amd.Reduced(g)
    D = O.weighted_farness(r.np, r.prp, r.pcol, r.pw)
    assert np.array_equal(r.farness_from(D), O.farness(g))


def test_raw_edgelist_roundtrip(tmp_path):
    g = bc_amd.buildGraph(O.dataset_path(RAW), raw=True)
    p = str(tmp_path / "m.txt")
    bc_amd.write_edgelist(p, g)
    h = bc_amd.buildGraph(p, raw=True)
    assert np.array_equal(h.row_ptr, g.row_ptr) and np.array_equal(h.col, g.col)
