"""Host-side library: CSR construction, loader, generators, exported symbols."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import ROOT


def test_csr_cleans_and_symmetrises():
    g = bc_amd.CSR.from_edges(5, [0, 1, 1, 2, 3, 3], [1, 0, 2, 2, 4, 4])
    assert g.n == 5 and g.nnz == 6
    assert g.row_ptr.tolist() == [0, 1, 3, 4, 5, 6]
    assert g.col.tolist() == [1, 0, 2, 1, 4, 3]
    assert g.degree.tolist() == [1, 2, 1, 1, 1]


def test_csr_lcc_relabels_in_order():
    g = bc_amd.CSR.from_edges(8, [0, 2, 3, 5, 6], [1, 3, 4, 6, 7], lcc=True)
    assert g.n == 3
    assert g.old_id.tolist() == [2, 3, 4]


def test_csr_rejects_bad_ids():
    with pytest.raises(bc_amd.CentError) as e:
        bc_amd.CSR.from_edges(3, [0], [5])
    assert e.value.status == bc_amd.CENT_E_INVAL


def test_loader_bases():
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    assert (g.startNodeID, g.endNodeID, g.n, g.m) == (1, 34, 35, 78)
    assert g.degree[0] == 0
    g = bc_amd.buildGraph(O.dataset_path("testdata2.txt"))
    assert g.startNodeID == 1 and g.m == 9
    g = bc_amd.buildGraph(O.dataset_path("ia-fb-messages.mtx"))
    assert g.startNodeID in (0, 1) and g.m > 0
    with pytest.raises(bc_amd.CentError) as e:
        bc_amd.buildGraph("/nonexistent/graph.txt")
    assert e.value.status == bc_amd.CENT_E_IO


def test_rmat_deterministic_and_sized():
    a = bc_amd.gen_rmat(12, 8 << 12, seed=1)
    b = bc_amd.gen_rmat(12, 8 << 12, seed=1)
    c = bc_amd.gen_rmat(12, 8 << 12, seed=2)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)
    assert not (a.n == c.n and np.array_equal(a.col[:100], c.col[:100]))
    # connected (LCC), no self loops, symmetric
    far = O.farness(a, [0])
    assert far[0] > 0
    u = np.repeat(np.arange(a.n), np.diff(a.row_ptr))
    assert not np.any(u == a.col)
    key = set(zip(u.tolist(), a.col.tolist()))
    assert all((y, x) in key for x, y in list(key)[:2000])


def test_grid_and_blocktree_generators():
    g = bc_amd.gen_grid(32, 32, 0.05, 10, seed=4)
    assert 900 < g.n <= 1024 and g.maxDegree <= 4 + 10
    t = bc_amd.gen_blocktree(50, 1.5, 8, 500, 6.0, 0.3, 0.5, seed=3)
    far = O.farness(t, [0, t.n - 1])
    assert np.all(far > 0)
    assert t.degree.min() >= 1
    assert (t.degree == 1).sum() > 0


def test_sample_sources():
    s = bc_amd.sample_sources(1000, 100, 7)
    assert len(set(s.tolist())) == 100 and s.min() >= 0 and s.max() < 1000
    assert np.array_equal(s, bc_amd.sample_sources(1000, 100, 7))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "cent.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", hdr))
    names -= {"if", "while", "for", "return", "sizeof"}
    assert "cent_bc" in names and "computeBC" in names
    lib = C.CDLL(bc_amd.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
