"""ctypes access to the CPU oracle (oracle/liboracle.so) — test infrastructure.

The oracle is the checker for the HIP engine; it is loaded only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes as C
import gzip
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_CC = os.path.join(ORACLE_DIR, "_ref", "ref_cc")
REF_BC = os.path.join(ORACLE_DIR, "_ref", "ref_bc")
DATASETS = os.path.join(ROOT, "tests", "golden", "datasets")


def _load():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    lib = C.CDLL(ORACLE_SO)
    I32P, I64P, F64P, F32P = (C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                              C.POINTER(C.c_double), C.POINTER(C.c_float))
    lib.oracle_bfs_farness.argtypes = [C.c_int32, I32P, I32P, I32P, C.c_int32, I64P, I64P]
    lib.oracle_brandes.argtypes = [C.c_int32, I32P, I32P, I32P, C.c_int32, C.c_int, C.c_int,
                                   F64P, I64P, I64P, I64P]
    lib.oracle_brandes_f32.argtypes = [C.c_int32, I32P, I32P, I32P, C.c_int32, F32P]
    lib.oracle_max_threads.restype = C.c_int
    lib.oracle_weighted_farness.argtypes = [C.c_int32, I32P, I32P, I64P, I32P, C.c_int32, I64P]
    lib.oracle_weighted_brandes.argtypes = [C.c_int32, I32P, I32P, I64P, I32P, C.c_int32, F64P, I64P]
    lib.oracle_weighted_brandes_src.argtypes = [C.c_int32, I32P, I32P, I64P, I32P, I64P, C.c_int32, F64P, I64P]
    return lib


_lib = _load()


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _arrays(g):
    rp = np.ascontiguousarray(g.row_ptr, np.int32)
    col = np.ascontiguousarray(g.col, np.int32) if len(g.col) else np.zeros(1, np.int32)
    return rp, col


def farness(g, sources=None, dist_sum=False):
    rp, col = _arrays(g)
    src = np.arange(g.n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    far = np.zeros(max(len(src), 1), np.int64)
    ds = np.zeros(max(g.n, 1), np.int64)
    st = _lib.oracle_bfs_farness(g.n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(src, C.c_int32),
                                 len(src), _p(far, C.c_int64), _p(ds, C.c_int64))
    assert st == 0
    return (far[:len(src)], ds[:g.n]) if dist_sum else far[:len(src)]


def brandes(g, sources=None, mode=0, threads=1):
    """Returns (bc, farness, dist_sum, dag_edges)."""
    rp, col = _arrays(g)
    src = np.arange(g.n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    bc = np.zeros(max(g.n, 1), np.float64)
    far = np.zeros(max(len(src), 1), np.int64)
    ds = np.zeros(max(g.n, 1), np.int64)
    dag = C.c_int64(0)
    st = _lib.oracle_brandes(g.n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(src, C.c_int32),
                             len(src), mode, threads, _p(bc, C.c_double), _p(far, C.c_int64),
                             _p(ds, C.c_int64), C.byref(dag))
    assert st == 0
    return bc[:g.n], far[:len(src)], ds[:g.n], dag.value


def brandes_f32(g, sources):
    rp, col = _arrays(g)
    src = np.ascontiguousarray(sources, np.int32)
    bc = np.zeros(max(g.n, 1), np.float32)
    assert _lib.oracle_brandes_f32(g.n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(src, C.c_int32),
                                   len(src), _p(bc, C.c_float)) == 0
    return bc[:g.n]


def weighted_farness(n, rp, col, w, sources=None):
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
    w = np.ascontiguousarray(w, np.int64) if len(w) else np.zeros(1, np.int64)
    src = np.arange(n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    far = np.zeros(max(len(src), 1), np.int64)
    assert _lib.oracle_weighted_farness(n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(w, C.c_int64),
                                        _p(src, C.c_int32), len(src), _p(far, C.c_int64)) == 0
    return far[:len(src)]


def weighted_brandes(n, rp, col, w, sources=None):
    """(bc, weighted farness) with source/target weights w."""
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
    w = np.ascontiguousarray(w, np.int64) if len(w) else np.zeros(1, np.int64)
    src = np.arange(n, dtype=np.int32) if sources is None else np.ascontiguousarray(sources, np.int32)
    bc = np.zeros(max(n, 1))
    far = np.zeros(max(len(src), 1), np.int64)
    assert _lib.oracle_weighted_brandes(n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(w, C.c_int64),
                                        _p(src, C.c_int32), len(src), _p(bc, C.c_double),
                                        _p(far, C.c_int64)) == 0
    return bc[:n], far[:len(src)]


def weighted_brandes_src(n, rp, col, w, sources, source_weights):
    """(bc, weighted farness) from listed sources with their own weights."""
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
    w = np.ascontiguousarray(w, np.int64) if len(w) else np.zeros(1, np.int64)
    src = np.ascontiguousarray(sources, np.int32) if len(sources) else np.zeros(1, np.int32)
    sw = np.ascontiguousarray(source_weights, np.int64) if len(sources) else np.zeros(1, np.int64)
    bc = np.zeros(max(n, 1))
    far = np.zeros(max(len(sources), 1), np.int64)
    assert _lib.oracle_weighted_brandes_src(n, _p(rp, C.c_int32), _p(col, C.c_int32), _p(w, C.c_int64),
                                            _p(src, C.c_int32), _p(sw, C.c_int64), len(sources),
                                            _p(bc, C.c_double), _p(far, C.c_int64)) == 0
    return bc[:n], far[:len(sources)]


def max_threads():
    return _lib.oracle_max_threads()


def dataset_path(name):
    """Path of a vendored dataset; .gz fixtures are unpacked to /tmp once."""
    p = os.path.join(DATASETS, name)
    if os.path.exists(p):
        return p
    gz = p + ".gz"
    out = os.path.join("/tmp", "bc_amd_" + name)
    if not os.path.exists(out):
        with gzip.open(gz, "rb") as f, open(out + ".part", "wb") as o:
            o.write(f.read())
        os.replace(out + ".part", out)
    return out
