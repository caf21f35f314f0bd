"""The reference's result-file formats (host only, no GPU)."""
import os

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN

SCORE = os.path.join(GOLDEN, "karate_BC_Score.txt")


def test_score_file_reads_reference_file():
    start, end, bc = bc_amd.read_bc_scores(SCORE)
    assert (start, end) == (1, 34)
    assert bc[1] == pytest.approx(792.491028) and bc[34] == pytest.approx(float(open(SCORE).read().split()[-1]))


def test_score_file_roundtrip_is_byte_identical(tmp_path):
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    _, _, bc = bc_amd.read_bc_scores(SCORE)
    out = str(tmp_path / "BC_Score.txt")
    bc_amd.write_bc_scores(out, g, bc)
    assert open(out, "rb").read() == open(SCORE, "rb").read()


def test_score_file_from_engine_formula_matches_reference(tmp_path):
    """Scores computed by the FP32 per-predecessor oracle, written in the
    reference format, agree with the reference's file line by line to 2e-6."""
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    f32 = O.brandes_f32(g, g.sources()).astype(np.float64)
    out = str(tmp_path / "BC_Score.txt")
    bc_amd.write_bc_scores(out, g, f32)
    got, ref = open(out).read().split(), open(SCORE).read().split()
    assert got[:2] == ref[:2] and len(got) == len(ref)
    np.testing.assert_allclose(np.array(got[2:], float), np.array(ref[2:], float), rtol=2e-6, atol=2e-6)


def test_ans_path_rule():
    assert bc_amd.ans_path("../dataset/ca-HepPh.mtx") == "../dataset/ca-HepPh_ans.txt"
    assert bc_amd.ans_path("tech-WHOIS.mtx") == "tech-WHOIS_ans.txt"
    with pytest.raises(bc_amd.CentError):
        bc_amd.ans_path("../dataset/karate.txt")


def test_answer_file_roundtrip(tmp_path):
    g = bc_amd.buildGraph(O.dataset_path("ia-fb-messages.mtx"))
    far = np.zeros(g.n, np.int64)
    far[g.sources()] = O.farness(g, g.sources())
    p = tmp_path / "ia-fb-messages.mtx"
    ans = bc_amd.ans_path(str(p))
    bc_amd.write_cc_answers(ans, g, far)
    lines = open(ans).read().splitlines()
    assert len(lines) == g.endNodeID - g.startNodeID + 1
    v, c = map(int, lines[0].split())
    assert v == g.startNodeID and c == far[v]
    got = bc_amd.read_cc_answers(ans, g)
    assert np.array_equal(got[g.sources()], far[g.sources()])


def test_answer_file_rejects_bad_input(tmp_path):
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    bad = tmp_path / "bad_ans.txt"
    bad.write_text("1 10\n2 11\n")  # too short: 34 lines expected
    with pytest.raises(bc_amd.CentError):
        bc_amd.read_cc_answers(str(bad), g)
    bad.write_text("".join(f"{v} 1\n" for v in range(1, 34)) + "99 3\n")  # id out of range
    with pytest.raises(bc_amd.CentError):
        bc_amd.read_cc_answers(str(bad), g)


def test_edgelist_writer_roundtrip(tmp_path):
    for g in (bc_amd.gen_rmat(10, 8 << 10, seed=3), bc_amd.buildGraph(O.dataset_path("dolphins.txt"))):
        p = str(tmp_path / "g.txt")
        bc_amd.write_edgelist(p, g)
        h = bc_amd.buildGraph(p)
        assert h.n == g.n and h.nnz == g.nnz
        assert np.array_equal(h.row_ptr, g.row_ptr) and np.array_equal(h.col, g.col)


def test_edgelist_writer_refuses_ambiguous_base(tmp_path):
    g = bc_amd.CSR.from_edges(4, np.array([1, 2]), np.array([2, 3]))  # 0-based, vertex 0 isolated
    with pytest.raises(bc_amd.CentError):
        bc_amd.write_edgelist(str(tmp_path / "x.txt"), g)
