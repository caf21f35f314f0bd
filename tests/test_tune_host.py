"""Autotuner table and graph signatures (host only, no GPU)."""
import os

import numpy as np
import pytest

import bc_amd


@pytest.fixture(autouse=True)
def clean_table():
    bc_amd.tune_table_clear()
    yield
    bc_amd.tune_table_clear()


def path_graph(n):
    u = np.arange(n - 1, dtype=np.int32)
    return bc_amd.CSR.from_edges(n, u, u + 1)


def star_graph(n):
    return bc_amd.CSR.from_edges(n, np.zeros(n - 1, np.int32), np.arange(1, n, dtype=np.int32))


def test_signature_closed_forms():
    # path of 16: avg degree 30/16 -> round(2 log2 1.875) = 2; max/avg = 2/1.875 -> 0;
    # depth 15 from an end -> log2(16) = 4
    assert bc_amd.tune_signature(path_graph(16), 15) == (4, 2, 0, 4)
    # star of 1025: avg degree 2048/1025 ~ 2 -> 2; skew log2(1024 / 2) = 9; depth 1 -> 1
    assert bc_amd.tune_signature(star_graph(1025), 1) == (10, 2, 9, 1)
    # R-MAT 2^14, 8 samples per vertex: average degree between 8 and 16 after dedup
    g = bc_amd.gen_rmat(14, 8 << 14, seed=1)
    lg_n, deg_x2, skew, lg_ecc = bc_amd.tune_signature(g, 6)
    assert lg_n == round(np.log2(g.n))
    assert deg_x2 == round(2 * np.log2(g.nnz / g.n))
    assert skew >= 5 and lg_ecc == 3


def test_table_exact_nearest_and_miss():
    a = bc_amd.tune_params(hub_deg=512, chunk_edges=2048, grid_waves=16384, frontier="dense", ms=1.5)
    b = bc_amd.tune_params(hub_deg=128, chunk_edges=512, frontier="sparse", ms=9.0)
    bc_amd.tune_table_put((20, 8, 12, 3), a)
    bc_amd.tune_table_put((22, 4, 2, 9), b)
    assert bc_amd.tune_table_size() == 2
    got = bc_amd.tune_table_get((20, 8, 12, 3))
    assert (got.hub_deg, got.chunk_edges, got.grid_waves, got.frontier) == (512, 2048, 16384, 1)
    near = bc_amd.tune_table_get((21, 8, 11, 3))          # distance 2 -> a
    assert near.hub_deg == 512
    assert bc_amd.tune_table_get((20, 8, 12, 6)) is None  # distance 3 -> miss
    assert bc_amd.tune_table_get((22, 4, 2, 8)).frontier == bc_amd.FRONTIER["sparse"]
    # same signature replaces
    bc_amd.tune_table_put((20, 8, 12, 3), b)
    assert bc_amd.tune_table_size() == 2
    assert bc_amd.tune_table_get((20, 8, 12, 3)).hub_deg == 128


def test_table_save_load_roundtrip(tmp_path):
    p = bc_amd.tune_params(hub_deg=1024, chunk_edges=4096, grid_waves=32768, frontier="dense", ms=3.25, push_pct=-1)
    q = bc_amd.tune_params(hub_deg=512, chunk_edges=512, grid_waves=0, frontier="source", ms=1.5)
    bc_amd.tune_table_put((10, 6, 5, 2), p, bc_amd.CENT_TUNE_BC)
    bc_amd.tune_table_put((10, 6, 5, 2), q, bc_amd.CENT_TUNE_CC)
    assert bc_amd.tune_table_size() == 2
    f = str(tmp_path / "t.txt")
    bc_amd.tune_table_save(f)
    bc_amd.tune_table_clear()
    assert bc_amd.tune_table_size() == 0
    bc_amd.tune_table_load(f)
    got = bc_amd.tune_table_get((10, 6, 5, 2))
    assert got.as_dict() == {"hub_deg": 1024, "chunk_edges": 4096, "grid_waves": 32768,
                             "frontier": "dense", "push_pct": -1, "ms": 3.25}
    assert bc_amd.tune_table_get((10, 6, 5, 2), bc_amd.CENT_TUNE_CC).frontier == bc_amd.FRONTIER["source"]


def test_table_entries_are_per_call_type(tmp_path):
    bc_amd.tune_table_put((12, 8, 6, 2), bc_amd.tune_params(frontier="source"), bc_amd.CENT_TUNE_BC)
    assert bc_amd.tune_table_get((12, 8, 6, 2), bc_amd.CENT_TUNE_CC) is None
    # a line without the call-type field is a betweenness entry
    f = tmp_path / "legacy.txt"
    f.write_text("# old format\n7 7 7 7  256 1024 0 1  2.0\n")
    bc_amd.tune_table_load(str(f))
    assert bc_amd.tune_table_get((7, 7, 7, 7), bc_amd.CENT_TUNE_BC).frontier == 1
    assert bc_amd.tune_table_get((7, 7, 7, 7), bc_amd.CENT_TUNE_CC) is None
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_put((1, 1, 1, 1), bc_amd.tune_params(), 3)


def test_table_rejects_bad_input(tmp_path):
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_put((1, 1, 1, 1), bc_amd.tune_params(hub_deg=8))
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_put((1, 1, 1, 1), bc_amd.TuneParams(256, 1024, 0, 7, 0, 0.0))
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_put((1, 1, 1, 1), bc_amd.tune_params(push_pct=-2))
    good = bc_amd.tune_params()
    bc_amd.tune_table_put((3, 3, 3, 3), good)
    f = tmp_path / "bad.txt"
    f.write_text("# header\n5 5 5 5 1  256 1024 0 1  1.0\n5 5 5\n")
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_load(str(f))
    assert bc_amd.tune_table_size() == 1        # nothing merged from the bad file
    with pytest.raises(bc_amd.CentError):
        bc_amd.tune_table_load(str(tmp_path / "absent.txt"))


def test_table_reads_lines_without_push_pct(tmp_path):
    f = tmp_path / "old.txt"
    f.write_text("# older form: no push_pct column\n12 10 6 2 2  4096 256 0 1  0.6745\n"
                 "12 10 6 2 512 512 0 1 0.8\n")
    bc_amd.tune_table_load(str(f))
    assert bc_amd.tune_table_get((12, 10, 6, 2), bc_amd.CENT_TUNE_CC).push_pct == 0
    got = bc_amd.tune_table_get((12, 10, 6, 2), bc_amd.CENT_TUNE_BC)
    assert (got.hub_deg, got.push_pct) == (512, 0)


def test_bundled_table_parses():
    if not os.path.exists(bc_amd.TUNE_TABLE):
        pytest.skip("no bundled table")
    bc_amd.tune_table_load()
    assert bc_amd.tune_table_size() >= 1
