"""The C-ABI from C: the reference's two mains with their hot path bound to
libcent.so (tests/capi/*.c, the patch INTEGRATION.md shows), compiled with
gcc against include/cent.h and run on the reference's karate fixture.  Their
output must equal the reference's own results: the BC score file
(karate_BC_Score.txt, FP32 per-predecessor formula, to its printed rounding)
and the D1/AP closeness the reference computes (ref_golden.json, written by
the reference compiled from its sources)."""
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import bc_amd
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CAPI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "capi")
LIBDIR = os.path.dirname(bc_amd.LIB_PATH)


def _build(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run(["gcc", "-O2", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(CAPI, name + ".c"), "-L", LIBDIR, "-lcent", f"-Wl,-rpath,{LIBDIR}",
                    "-o", exe], check=True)
    return exe


def test_c_brandes_main_reproduces_reference_scores(tmp_path):
    exe = _build(tmp_path, "brandes_main")
    out = str(tmp_path / "BC_Score.txt")
    r = subprocess.run([exe, O.dataset_path("karate.txt"), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, ref = open(out).read().split(), open(os.path.join(GOLDEN, "karate_BC_Score.txt")).read().split()
    assert got[:2] == ref[:2] and len(got) == len(ref)
    np.testing.assert_allclose(np.array(got[2:], float), np.array(ref[2:], float), rtol=2e-6, atol=2e-6)
    gold = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))["karate.txt"]
    np.testing.assert_allclose(np.array(got[2:], float), np.array(gold["bc_f32"], float), rtol=2e-6, atol=2e-6)
    # textbook formula through the same binary equals the oracle
    r = subprocess.run([exe, O.dataset_path("karate.txt"), out, "textbook"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    g = bc_amd.buildGraph(O.dataset_path("karate.txt"))
    want = O.brandes(g, g.sources())[0][g.sources()]
    np.testing.assert_allclose(np.array(open(out).read().split()[2:], float), want, rtol=1e-6, atol=1e-6)


def test_c_cc_main_reproduces_reference_closeness(tmp_path):
    exe = _build(tmp_path, "cc_main")
    r = subprocess.run([exe, O.dataset_path("karate.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.array([list(map(int, line.split())) for line in r.stdout.splitlines()])
    gold = json.load(open(os.path.join(GOLDEN, "ref_golden.json")))["karate.txt"]
    assert rows[:, 0].tolist() == list(range(1, 35))
    assert rows[:, 1].tolist() == gold["d1ap"]
    assert rows[:, 2].tolist() == gold["cc"]
    assert rows[6, 1] == 86  # CCs[7], Lib/AP_Process/AP_process_karate.txt
