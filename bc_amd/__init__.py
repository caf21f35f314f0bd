"""bc_amd — MI355X-native betweenness / closeness engine (Python host mirror).

A thin ctypes layer over the C-ABI in ``include/cent.h`` (``bc_amd/libcent.so``).
The names mirror the reference's operator interface for the hot path:

=====================================  ======================================================
``buildGraph(path)`` / ``createCSR``   Lib/AdjList/AdjList.c:13-150, Lib/CSR/CSR.c:9-108
``computeBC(csr, BCs)``                Seq_Brandes/Brandes.c:224-372 (BCs accumulated, +=)
``computeCC(csr, CCs)``                Seq_CC/CC.cpp:37-117 (CCs assigned per source id)
``compute_D1_AP_CC_shareBased_DegreeOrder(csr, CCs)``  Seq_CC/CC.cpp:1728-2052
=====================================  ======================================================

Everything runs on the HIP engine; there is no CPU fallback.  If the library
is not built this module raises ImportError, and engine calls on a machine
without a gfx950 device raise CentError(CENT_E_DEVICE).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcent.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `make -C bc_amd`")

_lib = C.CDLL(LIB_PATH)

# ---------------------------------------------------------------- statuses
CENT_OK, CENT_E_INVAL, CENT_E_OOM, CENT_E_DEVICE, CENT_E_OVERFLOW, CENT_E_IO = 0, 1, 2, 3, 4, 5
CENT_E_BUSY, CENT_E_CANCELLED, CENT_E_NOTFOUND, CENT_E_UNSUPPORTED = 6, 7, 8, 9
CENT_BC_TEXTBOOK, CENT_BC_REF_COMPAT = 0, 1
CENT_OPT_TIMING = 1
CENT_OPT_SOURCE_ORDER = 2
CENT_OPT_ONE_STREAM = 4
CENT_CSR_LCC = 1
CENT_CSR_RAW = 2
CENT_CSR_KIND_UNKNOWN, CENT_CSR_KIND_SIMPLE, CENT_CSR_KIND_MULTI = 0, 1, 2

_lib.cent_strerror.restype = C.c_char_p
_lib.cent_strerror.argtypes = [C.c_int]


class CentError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        msg = f"{where}: {_lib.cent_strerror(status).decode()} (status {status})"
        if detail:
            msg += f" [{detail}]"
        super().__init__(msg)


def _check(status: int, where: str, detail: str = "") -> None:
    if status != CENT_OK:
        raise CentError(status, where, detail)


# ---------------------------------------------------------------- CSR
class _CSR(C.Structure):
    _fields_ = [
        ("csrV", C.POINTER(C.c_int32)),
        ("csrE", C.POINTER(C.c_int32)),
        ("csrNodesDegree", C.POINTER(C.c_int32)),
        ("csrESize", C.c_int64),
        ("csrVSize", C.c_int32),
        ("startNodeID", C.c_int32),
        ("endNodeID", C.c_int32),
        ("totalNodeNumber", C.c_int32),
        ("maxDegree", C.c_int32),
        ("startAtZero", C.c_int32),
        ("kind", C.c_int32),
    ]


_P_CSR = C.POINTER(_CSR)
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_F64P = C.POINTER(C.c_double)

_lib.cent_csr_from_edges.argtypes = [C.c_int32, _I32P, _I32P, C.c_int64, C.c_int32,
                                     C.POINTER(_P_CSR), C.POINTER(_I32P)]
_lib.cent_csr_from_arrays.argtypes = [C.c_int32, _I32P, _I32P, C.POINTER(_P_CSR)]
_lib.cent_csr_free.argtypes = [_P_CSR]
_lib.cent_csr_free.restype = None
_lib.cent_free.argtypes = [C.c_void_p]
_lib.cent_free.restype = None
_lib.cent_load_edgelist.argtypes = [C.c_char_p, C.POINTER(_P_CSR)]
_lib.cent_load_edgelist_ex.argtypes = [C.c_char_p, C.c_int32, C.POINTER(_P_CSR)]
_lib.cent_gen_rmat.argtypes = [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double,
                               C.c_uint64, C.c_int32, C.POINTER(_P_CSR)]
_lib.cent_gen_grid.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_uint64,
                               C.POINTER(_P_CSR)]
_lib.cent_gen_blocktree.argtypes = [C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_double,
                                    C.c_double, C.c_double, C.c_uint64, C.POINTER(_P_CSR)]
_lib.cent_sample_sources.argtypes = [C.c_int32, C.c_int32, C.c_uint64, _I32P]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class CSR:
    """Owned handle on a library ``struct CSR`` (0-based, immutable).

    ``row_ptr``/``col`` are zero-copy numpy views of the library's arrays.
    """

    def __init__(self, ptr: _P_CSR):
        if not ptr:
            raise ValueError("null CSR")
        self._p = ptr
        s = ptr.contents
        self.n = int(s.csrVSize) - 1
        self.nnz = int(s.csrESize)
        self.startNodeID = int(s.startNodeID)
        self.endNodeID = int(s.endNodeID)
        self.totalNodeNumber = int(s.totalNodeNumber)
        self.maxDegree = int(s.maxDegree)
        self.startAtZero = int(s.startAtZero)
        self.kind = int(s.kind)  # CENT_CSR_KIND_*: simple graph, multigraph or not known
        self.row_ptr = np.ctypeslib.as_array(s.csrV, shape=(self.n + 1,))
        self.col = (np.ctypeslib.as_array(s.csrE, shape=(self.nnz,)) if self.nnz
                    else np.zeros(0, np.int32))
        self.degree = (np.ctypeslib.as_array(s.csrNodesDegree, shape=(self.n,)) if self.n
                       else np.zeros(0, np.int32))

    @property
    def m(self) -> int:
        """Undirected edge count."""
        return self.nnz // 2

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            _lib.cent_csr_free(p)
            self._p = None

    @classmethod
    def from_edges(cls, n: int, u, v, lcc: bool = False, raw: bool = False):
        """raw=True keeps repeated edges and self-loops (CENT_CSR_RAW)."""
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        if u.shape != v.shape:
            raise ValueError("u and v differ in length")
        out = _P_CSR()
        old = _I32P()
        _check(_lib.cent_csr_from_edges(n, _ptr(u, C.c_int32), _ptr(v, C.c_int32), len(u),
                                        (CENT_CSR_LCC if lcc else 0) | (CENT_CSR_RAW if raw else 0), C.byref(out),
                                        C.byref(old) if lcc else None), "cent_csr_from_edges")
        g = cls(out)
        if lcc and old:
            g.old_id = np.ctypeslib.as_array(old, shape=(g.n,)).copy() if g.n else np.zeros(0, np.int32)
            _lib.cent_free(old)
        return g

    @classmethod
    def from_arrays(cls, row_ptr, col):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cl = np.ascontiguousarray(col, dtype=np.int32)
        out = _P_CSR()
        _check(_lib.cent_csr_from_arrays(len(rp) - 1, _ptr(rp, C.c_int32), _ptr(cl, C.c_int32),
                                         C.byref(out)), "cent_csr_from_arrays")
        return cls(out)

    def sources(self) -> np.ndarray:
        """The reference's source range startNodeID..endNodeID."""
        return np.arange(self.startNodeID, self.endNodeID + 1, dtype=np.int32)


def buildGraph(path: str, raw: bool = False) -> CSR:
    """Load a reference-format edge list (Lib/AdjList/AdjList.c:13-150).
    raw=True keeps repeated edges and self-loops as the reference's loader
    does (CENT_CSR_RAW); by default they are removed."""
    out = _P_CSR()
    _check(_lib.cent_load_edgelist_ex(os.fsencode(path), CENT_CSR_RAW if raw else 0, C.byref(out)),
           "buildGraph", path)
    return CSR(out)


def createCSR(graph: CSR) -> CSR:
    """The loader already yields a CSR (Lib/CSR/CSR.c:9); kept for call-site parity."""
    return graph


def gen_rmat(scale: int, samples: int, a=0.57, b=0.19, c=0.19, seed=1, lcc=True) -> CSR:
    out = _P_CSR()
    _check(_lib.cent_gen_rmat(scale, samples, a, b, c, seed, CENT_CSR_LCC if lcc else 0,
                              C.byref(out)), "cent_gen_rmat")
    return CSR(out)


def gen_grid(rows: int, cols: int, del_frac=0.05, shortcuts=1000, seed=4) -> CSR:
    out = _P_CSR()
    _check(_lib.cent_gen_grid(rows, cols, del_frac, shortcuts, seed, C.byref(out)), "cent_gen_grid")
    return CSR(out)


def gen_blocktree(n_blocks=2000, alpha=1.5, smin=8, smax=50000, avg_deg=6.0,
                  pendant_frac=0.3, p_geo=0.5, seed=3) -> CSR:
    out = _P_CSR()
    _check(_lib.cent_gen_blocktree(n_blocks, alpha, smin, smax, avg_deg, pendant_frac, p_geo,
                                   seed, C.byref(out)), "cent_gen_blocktree")
    return CSR(out)


def sample_sources(n: int, k: int, seed: int) -> np.ndarray:
    out = np.empty(k, np.int32)
    _check(_lib.cent_sample_sources(n, k, seed, _ptr(out, C.c_int32)), "cent_sample_sources")
    return out


# ---------------------------------------------------------------- reduced form
class _RView(C.Structure):
    _fields_ = [("n", C.c_int32), ("np", C.c_int32), ("nnzp", C.c_int64), ("nblocks", C.c_int32),
                ("nitems", C.c_int32), ("nfold", C.c_int32), ("n_ap", C.c_int32),
                ("ncomp", C.c_int32), ("prp", _I32P), ("pcol", _I32P), ("pw", _I64P),
                ("porig", _I32P), ("boff", _I32P)]


_lib.cent_reduce.argtypes = [_P_CSR, C.POINTER(C.c_void_p)]
_lib.cent_reduce_gpu.argtypes = [C.c_void_p, _P_CSR, C.POINTER(C.c_void_p)]
_lib.cent_reduced_get_view.argtypes = [C.c_void_p, C.POINTER(_RView)]
_lib.cent_reduced_farness.argtypes = [C.c_void_p, _I64P, _I64P]
_lib.cent_reduced_bc.argtypes = [C.c_void_p, _F64P, _F64P]
_lib.cent_reduced_free.argtypes = [C.c_void_p]
_lib.cent_reduced_source_weights.argtypes = [C.c_void_p, _I32P, C.c_int32, _I64P]
_lib.cent_reduced_bc_sources.argtypes = [C.c_void_p, _I32P, C.c_int32, _F64P, _F64P]
_lib.cent_reduced_farness_sources.argtypes = [C.c_void_p, _I64P, _I32P, C.c_int32, _I64P]
_lib.cent_reduced_free.restype = None
CENT_SCHED_PERSISTENT, CENT_SCHED_PER_ITEM, CENT_SCHED_SHARED = 0, 1, 2  # see include/cent.h


class Reduced:
    """Degree-1 folded, biconnected-block packed form of a graph.

    Mirrors the reduction steps of compute_D1_AP_CC_shareBased_DegreeOrder
    (Seq_CC/CC.cpp:1742-1776: D1Folding, AP_detection + AP_Copy_And_Split,
    rebuildGraph).  Built by the host C library (cent_reduce), or, with
    `ctx`, by GPU kernels on that context's device (cent_reduce_gpu); both
    give the same blocks, weights and results.  Arrays are zero-copy views
    of the library's buffers.
    """

    def __init__(self, csr: CSR, ctx: "Context" = None):
        self._csr = csr
        self._ctx = ctx
        h = C.c_void_p()
        if ctx is None:
            _check(_lib.cent_reduce(csr._p, C.byref(h)), "cent_reduce")
        else:
            _check(_lib.cent_reduce_gpu(ctx._h, csr._p, C.byref(h)), "cent_reduce_gpu", ctx.last_error())
        self._h = h
        v = _RView()
        _check(_lib.cent_reduced_get_view(h, C.byref(v)), "cent_reduced_get_view")
        self.n, self.np, self.nnzp = v.n, v.np, v.nnzp
        self.nblocks, self.nitems, self.nfold, self.n_ap, self.ncomp = (
            v.nblocks, v.nitems, v.nfold, v.n_ap, v.ncomp)
        arr = np.ctypeslib.as_array
        self.prp = arr(v.prp, shape=(self.np + 1,)) if v.prp else np.zeros(1, np.int32)
        self.pcol = arr(v.pcol, shape=(self.nnzp,)) if self.nnzp else np.zeros(0, np.int32)
        self.pw = arr(v.pw, shape=(self.np,)) if self.np else np.zeros(0, np.int64)
        self.porig = arr(v.porig, shape=(self.np,)) if self.np else np.zeros(0, np.int32)
        self.boff = arr(v.boff, shape=(self.nblocks + 1,)) if v.boff else np.zeros(1, np.int32)

    def farness_from(self, D) -> np.ndarray:
        """farness of every input vertex from the packed weighted distance sums D."""
        D = np.ascontiguousarray(D, dtype=np.int64)
        if len(D) != self.np:
            raise ValueError("D must have one entry per packed vertex")
        out = np.zeros(max(self.n, 1), np.int64)
        _check(_lib.cent_reduced_farness(self._h, _ptr(D, C.c_int64) if self.np else None,
                                         _ptr(out, C.c_int64)), "cent_reduced_farness")
        return out[:self.n]

    def workspace_estimate(self, what: int = None, k: int = -1, sched: int = CENT_SCHED_PERSISTENT,
                           one_stream: bool = False) -> int:
        """cent_reduced_workspace_estimate: device bytes a context allocates
        at most for cent_cc_reduced (CENT_TUNE_CC, the default),
        cent_bc_reduced (CENT_TUNE_BC, k = -1) or cent_bc_reduced_sources
        (k sources)."""
        what = CENT_TUNE_CC if what is None else what
        out = C.c_int64()
        o = _opts(one_stream=one_stream)
        _check(_lib.cent_reduced_workspace_estimate(self._h, what, k, sched, C.addressof(o), C.byref(out)),
               "cent_reduced_workspace_estimate")
        return int(out.value)

    def source_weights(self, sources) -> np.ndarray:
        """WS[x]: sources (with repeats) hanging off packed copy x (pw for all vertices once)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        out = np.zeros(max(self.np, 1), np.int64)
        _check(_lib.cent_reduced_source_weights(self._h, _ptr(src, C.c_int32) if len(src) else None, len(src),
                                                _ptr(out, C.c_int64)), "cent_reduced_source_weights")
        return out[:self.np]

    def bc_from_sources(self, sources, bcp) -> np.ndarray:
        """betweenness from the source multiset, given the packed dependencies bcp
        (weighted Brandes from every copy with source weights WS)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        bcp = np.ascontiguousarray(bcp, dtype=np.float64)
        out = np.zeros(max(self.n, 1))
        _check(_lib.cent_reduced_bc_sources(self._h, _ptr(src, C.c_int32) if len(src) else None, len(src),
                                            _ptr(bcp, C.c_double) if self.np else None, _ptr(out, C.c_double)),
               "cent_reduced_bc_sources")
        return out[:self.n]

    def farness_of_sources(self, sources, D) -> np.ndarray:
        src = np.ascontiguousarray(sources, dtype=np.int32)
        D = np.ascontiguousarray(D, dtype=np.int64)
        out = np.zeros(max(len(src), 1), np.int64)
        _check(_lib.cent_reduced_farness_sources(self._h, _ptr(D, C.c_int64) if self.np else None,
                                                 _ptr(src, C.c_int32) if len(src) else None, len(src),
                                                 _ptr(out, C.c_int64)), "cent_reduced_farness_sources")
        return out[:len(src)]

    def bc_from(self, bcp) -> np.ndarray:
        """betweenness of every input vertex from the packed weighted dependencies bcp."""
        bcp = np.ascontiguousarray(bcp, dtype=np.float64)
        if len(bcp) != self.np:
            raise ValueError("bcp must have one entry per packed vertex")
        out = np.zeros(max(self.n, 1))
        _check(_lib.cent_reduced_bc(self._h, _ptr(bcp, C.c_double) if self.np else None,
                                    _ptr(out, C.c_double)), "cent_reduced_bc")
        return out[:self.n]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.cent_reduced_free(h)
            self._h = None


# ---------------------------------------------------------------- engine
class _Opts(C.Structure):
    _fields_ = [("bc_mode", C.c_int32), ("flags", C.c_int32), ("block_size", C.c_int32),
                ("grid_waves", C.c_int32), ("frontier", C.c_int32), ("max_slots", C.c_int32),
                ("push_pct", C.c_int32)]


CENT_FRONTIER_AUTO, CENT_FRONTIER_DENSE, CENT_FRONTIER_SPARSE = 0, 1, 2
CENT_FRONTIER_SOURCE = 3
FRONTIER = {"auto": CENT_FRONTIER_AUTO, "dense": CENT_FRONTIER_DENSE, "sparse": CENT_FRONTIER_SPARSE,
            "source": CENT_FRONTIER_SOURCE}


class Stats(C.Structure):
    _fields_ = [("ms_fwd", C.c_double), ("ms_bwd", C.c_double), ("ms_total", C.c_double),
                ("launches_fwd", C.c_int64), ("launches_bwd", C.c_int64),
                ("batches", C.c_int64), ("levels", C.c_int64), ("dag_edges", C.c_int64),
                ("reached", C.c_int64), ("edges_reached", C.c_int64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib.cent_ctx_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
_lib.cent_ctx_destroy.argtypes = [C.c_void_p]
_lib.cent_ctx_destroy.restype = None
_lib.cent_ctx_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
_lib.cent_ctx_reset_stats.argtypes = [C.c_void_p]
_lib.cent_ctx_last_error.argtypes = [C.c_void_p]
_lib.cent_ctx_last_error.restype = C.c_char_p
_lib.cent_graph_from_csr.argtypes = [C.c_void_p, _I32P, _I32P, C.c_int32, C.c_int64,
                                     C.POINTER(C.c_void_p)]
_lib.cent_graph_create.argtypes = [C.c_void_p, _I32P, _I32P, C.c_int32, C.c_int64, C.c_int32,
                                   C.POINTER(C.c_void_p)]
CENT_GRAPH_KEEP_ORDER = 1
CENT_GRAPH_DEGREE_ORDER = 4
CENT_GRAPH_BFS_ORDER = 8
_ORDER_FLAGS = {"auto": 0, "keep": CENT_GRAPH_KEEP_ORDER, "degree": CENT_GRAPH_DEGREE_ORDER,
                "bfs": CENT_GRAPH_BFS_ORDER}
_lib.cent_graph_free.argtypes = [C.c_void_p]
_lib.cent_graph_free.restype = None
_lib.cent_bc.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, _F64P, _I64P, _I64P,
                         C.POINTER(_Opts)]
_lib.cent_cc.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, _I64P, _I64P,
                         C.POINTER(_Opts)]
_lib.cent_bc_device.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.POINTER(_Opts)]
_lib.cent_cc_device.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, C.c_void_p, C.c_void_p,
                                C.POINTER(_Opts)]
CENT_COPY_TO_DEVICE, CENT_COPY_TO_HOST, CENT_COPY_ON_DEVICE = 1, 2, 3
_lib.cent_cc_device_noalloc.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, C.c_void_p, C.c_void_p]
_lib.cent_device_alloc.argtypes = [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]
_lib.cent_device_free.argtypes = [C.c_void_p, C.c_void_p]
_lib.cent_device_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32]
_lib.cent_device_fill.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int64]
_lib.cent_ctx_workspace_bytes.argtypes = [C.c_void_p]
_lib.cent_ctx_workspace_bytes.restype = C.c_int64
_lib.cent_ctx_trim.argtypes = [C.c_void_p]
_lib.cent_probe_copy_bandwidth.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
_lib.cent_probe_gather_bandwidth.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
_lib.cent_probe_gather_lines.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
_lib.cent_cc_reduced.argtypes = [C.c_void_p, C.c_void_p, _I64P, C.c_int32, C.POINTER(_Opts)]
_lib.cent_bc_reduced.argtypes = [C.c_void_p, C.c_void_p, _F64P, _I64P, C.c_int32, C.POINTER(_Opts)]
_lib.cent_bc_reduced_sources.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, _F64P, _I64P, C.c_int32,
                                         C.POINTER(_Opts)]
_lib.compute_D1_AP_CC_shareBased_DegreeOrder.argtypes = [_P_CSR, _I64P]
_lib.computeBC.argtypes = [_P_CSR, _F64P]
_lib.computeBC_compat.argtypes = [_P_CSR, _F64P]
_lib.computeCC.argtypes = [_P_CSR, _I64P]


# ------------------------------------------------------- result files
_lib.cent_write_bc_scores.argtypes = [C.c_char_p, _P_CSR, _F64P]
_lib.cent_read_bc_scores.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _F64P, C.c_int32]
_lib.cent_ans_path.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
_lib.cent_read_cc_answers.argtypes = [C.c_char_p, _P_CSR, _I64P]
_lib.cent_write_cc_answers.argtypes = [C.c_char_p, _P_CSR, _I64P]


_lib.cent_write_edgelist.argtypes = [C.c_char_p, _P_CSR]


def write_edgelist(path: str, csr: "CSR") -> None:
    """csr as a reference edge-list file ("n m", then "u v" per undirected edge)."""
    _check(_lib.cent_write_edgelist(path.encode(), csr._p), "cent_write_edgelist", path)


def write_bc_scores(path: str, csr: "CSR", bc) -> None:
    """BC score file in the reference's format ("start end", then "%f" per vertex)."""
    b = np.ascontiguousarray(bc, dtype=np.float64)
    _check(_lib.cent_write_bc_scores(path.encode(), csr._p, _ptr(b, C.c_double)), "cent_write_bc_scores", path)


def read_bc_scores(path: str, cap: int = 1 << 26):
    """(start, end, scores) with scores[start..end] from a reference score file."""
    out = np.zeros(cap, np.float64)
    s, e = C.c_int32(), C.c_int32()
    _check(_lib.cent_read_bc_scores(path.encode(), C.byref(s), C.byref(e), _ptr(out, C.c_double), cap),
           "cent_read_bc_scores", path)
    return s.value, e.value, out[:e.value + 1].copy()


def ans_path(dataset_path: str) -> str:
    """The reference's answer-file name for a .mtx dataset."""
    buf = C.create_string_buffer(4096)
    _check(_lib.cent_ans_path(dataset_path.encode(), buf, 4096), "cent_ans_path", dataset_path)
    return buf.value.decode()


def read_cc_answers(path: str, csr: "CSR") -> np.ndarray:
    ans = np.zeros(max(csr.n, 1), np.int64)
    _check(_lib.cent_read_cc_answers(path.encode(), csr._p, _ptr(ans, C.c_int64)), "cent_read_cc_answers", path)
    return ans[:csr.n]


def write_cc_answers(path: str, csr: "CSR", cc) -> None:
    c = np.ascontiguousarray(cc, dtype=np.int64)
    _check(_lib.cent_write_cc_answers(path.encode(), csr._p, _ptr(c, C.c_int64)), "cent_write_cc_answers", path)


class TuneParams(C.Structure):
    """Launch settings of the plain form (``cent_tune_params``)."""
    _fields_ = [("hub_deg", C.c_int32), ("chunk_edges", C.c_int32), ("grid_waves", C.c_int32),
                ("frontier", C.c_int32), ("push_pct", C.c_int32), ("ms", C.c_double)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["frontier"] = {v: k for k, v in FRONTIER.items()}.get(self.frontier, self.frontier)
        return d


class TuneSig(C.Structure):
    """Graph signature the tuning table is keyed on (``cent_tune_sig``)."""
    _fields_ = [("lg_n", C.c_int32), ("deg_x2", C.c_int32), ("skew", C.c_int32), ("lg_ecc", C.c_int32)]

    def as_tuple(self) -> tuple:
        return (self.lg_n, self.deg_x2, self.skew, self.lg_ecc)


CENT_TUNE_BC = 1
CENT_TUNE_CC = 2
CENT_GRAPH_NO_TUNE = 2
_lib.cent_graph_get_layout.argtypes = [C.c_void_p, C.c_int32, C.POINTER(TuneParams)]
_lib.cent_graph_set_layout.argtypes = [C.c_void_p, C.c_int32, C.POINTER(TuneParams)]
_lib.cent_graph_signature.argtypes = [C.c_void_p, C.POINTER(TuneSig)]
_lib.cent_tune.argtypes = [C.c_void_p, C.c_void_p, _I32P, C.c_int32, C.c_int32, C.c_int32,
                           C.POINTER(TuneParams)]
_lib.cent_tune_signature.argtypes = [_I32P, C.c_int32, C.c_int64, C.c_int32, C.POINTER(TuneSig)]
_lib.cent_tune_params_check.argtypes = [C.POINTER(TuneParams)]
_lib.cent_tune_table_put.argtypes = [C.POINTER(TuneSig), C.c_int32, C.POINTER(TuneParams)]
_lib.cent_tune_table_get.argtypes = [C.POINTER(TuneSig), C.c_int32, C.POINTER(TuneParams)]
_lib.cent_tune_table_size.restype = C.c_int32
_lib.cent_tune_table_clear.restype = None
_lib.cent_tune_table_load.argtypes = [C.c_char_p]
_lib.cent_tune_table_save.argtypes = [C.c_char_p]

#: Tuning table measured on MI355X for the bundled workloads (loaded by
#: ``tune_table_load()`` with no argument).
TUNE_TABLE = os.path.join(_HERE, "tune_mi355x.txt")


def tune_signature(csr: "CSR", ecc: int) -> tuple:
    """Signature of a host CSR given its BFS-depth probe (host only)."""
    sig = TuneSig()
    _check(_lib.cent_tune_signature(_ptr(csr.row_ptr, C.c_int32), csr.n, csr.nnz, ecc, C.byref(sig)),
           "cent_tune_signature")
    return sig.as_tuple()


def tune_params(hub_deg=256, chunk_edges=1024, grid_waves=0, frontier="auto", ms=0.0, push_pct=0) -> TuneParams:
    return TuneParams(hub_deg, chunk_edges, grid_waves, FRONTIER[frontier], push_pct, ms)


def tune_table_put(sig: tuple, p: TuneParams, what: int = CENT_TUNE_BC) -> None:
    _check(_lib.cent_tune_table_put(C.byref(TuneSig(*sig)), what, C.byref(p)), "cent_tune_table_put")


def tune_table_get(sig: tuple, what: int = CENT_TUNE_BC):
    """Entry for call type ``what`` nearest to ``sig`` (exact, else within 2 buckets), or None."""
    p = TuneParams()
    st = _lib.cent_tune_table_get(C.byref(TuneSig(*sig)), what, C.byref(p))
    if st == 8:  # CENT_E_NOTFOUND
        return None
    _check(st, "cent_tune_table_get")
    return p


def tune_table_size() -> int:
    return int(_lib.cent_tune_table_size())


def tune_table_clear() -> None:
    _lib.cent_tune_table_clear()


def tune_table_load(path: str = None) -> None:
    _check(_lib.cent_tune_table_load((path or TUNE_TABLE).encode()), "cent_tune_table_load", path or TUNE_TABLE)


def tune_table_save(path: str) -> None:
    _check(_lib.cent_tune_table_save(path.encode()), "cent_tune_table_save", path)


def _opts(mode=CENT_BC_TEXTBOOK, timing=False, grid_waves=0, frontier="auto", max_slots=0,
          keep_source_order=False, one_stream=False, push_pct=0) -> _Opts:
    o = _Opts()
    o.bc_mode = mode
    o.flags = ((CENT_OPT_TIMING if timing else 0) | (CENT_OPT_SOURCE_ORDER if keep_source_order else 0) |
               (CENT_OPT_ONE_STREAM if one_stream else 0))
    o.grid_waves = grid_waves
    o.frontier = FRONTIER[frontier]
    o.max_slots = max_slots
    o.push_pct = push_pct
    return o


_lib.cent_workspace_estimate.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_void_p, C.POINTER(C.c_int64)]
_lib.cent_graph_workspace_estimate.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.POINTER(C.c_int64)]
_lib.cent_reduced_workspace_estimate.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                                 C.POINTER(C.c_int64)]


def workspace_estimate(csr: CSR, what: int = CENT_TUNE_BC, k: int = -1, depth: int = -1, tuned: bool = True,
                       mode: int = CENT_BC_TEXTBOOK, grid_waves: int = 0, frontier: str = "auto",
                       max_slots: int = 0, one_stream: bool = False) -> int:
    """Device bytes a context allocates at most for one plain call (cent_bc
    for CENT_TUNE_BC, cent_cc for CENT_TUNE_CC) with k sources (-1: all) on
    csr before it is uploaded (cent_workspace_estimate).  depth: BFS depth
    from the highest-degree vertex, -1 if not known."""
    out = C.c_int64()
    o = _opts(mode, False, grid_waves, frontier, max_slots, False, one_stream)
    _check(_lib.cent_workspace_estimate(csr.row_ptr.ctypes.data, csr.n, csr.nnz, depth,
                                        0 if tuned else CENT_GRAPH_NO_TUNE, what, k, C.addressof(o),
                                        C.byref(out)), "cent_workspace_estimate")
    return int(out.value)


_lib.cent_reduce_gpu_workspace_estimate.argtypes = [C.c_int32, C.c_int64, C.POINTER(C.c_int64)]


def reduce_gpu_workspace_estimate(csr: CSR) -> int:
    """Device scratch cent_reduce_gpu takes for csr (cent_reduce_gpu_workspace_estimate)."""
    out = C.c_int64()
    _check(_lib.cent_reduce_gpu_workspace_estimate(csr.n, csr.nnz, C.byref(out)),
           "cent_reduce_gpu_workspace_estimate")
    return int(out.value)


class DeviceArray:
    """n elements of a numpy dtype in device memory of a context's GPU
    (cent_device_alloc), freed with the object."""

    def __init__(self, ctx: "Context", n: int, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _check(_lib.cent_device_alloc(ctx._h, self.nbytes, C.byref(p)), "cent_device_alloc",
               ctx.last_error())
        self.ptr = p.value

    def zero(self) -> "DeviceArray":
        _check(_lib.cent_device_fill(self.ctx._h, self.ptr, 0, self.nbytes), "cent_device_fill",
               self.ctx.last_error())
        return self

    def copy_from(self, a) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"copy_from: {a.size} elements into {self.n}")
        _check(_lib.cent_device_copy(self.ctx._h, self.ptr, a.ctypes.data, self.nbytes,
                                     CENT_COPY_TO_DEVICE), "cent_device_copy", self.ctx.last_error())
        return self

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        _check(_lib.cent_device_copy(self.ctx._h, out.ctypes.data, self.ptr, self.nbytes,
                                     CENT_COPY_TO_HOST), "cent_device_copy", self.ctx.last_error())
        return out

    def __del__(self):
        p, self.ptr = getattr(self, "ptr", None), None
        if p and getattr(self.ctx, "_h", None):
            _lib.cent_device_free(self.ctx._h, p)


class DeviceGraph:
    def __init__(self, ctx: "Context", csr: CSR, keep_order: bool = False, tuned: bool = True,
                 order: str = "auto"):
        """order: device relabel, "auto" (degree order, or BFS order for graphs
        deeper than 48 levels from the highest-degree vertex), "degree", "bfs"
        or "keep" (the caller's ids; the same as keep_order=True)."""
        self.ctx = ctx
        self.n = csr.n
        self.nnz = csr.nnz
        self._csr = csr
        h = C.c_void_p()
        if order not in _ORDER_FLAGS:
            raise ValueError(f"order must be one of {sorted(_ORDER_FLAGS)}")
        flags = ((CENT_GRAPH_KEEP_ORDER if keep_order else _ORDER_FLAGS[order])
                 | (0 if tuned else CENT_GRAPH_NO_TUNE))
        _check(_lib.cent_graph_create(ctx._h, _ptr(csr.row_ptr, C.c_int32),
                                      _ptr(csr.col, C.c_int32) if csr.nnz else None,
                                      csr.n, csr.nnz, flags, C.byref(h)),
               "cent_graph_create", ctx.last_error())
        self._h = h

    def layout_for(self, what: int) -> TuneParams:
        """Launch settings used for betweenness (CENT_TUNE_BC) or closeness (CENT_TUNE_CC) calls."""
        p = TuneParams()
        _check(_lib.cent_graph_get_layout(self._h, what, C.byref(p)), "cent_graph_get_layout")
        return p

    @property
    def layout(self) -> TuneParams:
        return self.layout_for(CENT_TUNE_BC)

    def set_layout(self, p: TuneParams, what: int = CENT_TUNE_BC | CENT_TUNE_CC) -> None:
        _check(_lib.cent_graph_set_layout(self._h, what, C.byref(p)), "cent_graph_set_layout",
               self.ctx.last_error())

    def workspace_estimate(self, what: int = CENT_TUNE_BC, k: int = -1, mode: int = CENT_BC_TEXTBOOK,
                           grid_waves: int = 0, frontier: str = "auto", max_slots: int = 0,
                           one_stream: bool = False) -> int:
        """cent_graph_workspace_estimate: device bytes a context allocates at
        most for one plain call on this graph."""
        out = C.c_int64()
        o = _opts(mode, False, grid_waves, frontier, max_slots, False, one_stream)
        _check(_lib.cent_graph_workspace_estimate(self._h, what, k, C.addressof(o), C.byref(out)),
               "cent_graph_workspace_estimate")
        return int(out.value)

    @property
    def signature(self) -> tuple:
        sig = TuneSig()
        _check(_lib.cent_graph_signature(self._h, C.byref(sig)), "cent_graph_signature")
        return sig.as_tuple()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.cent_graph_free(h)
            self._h = None


class Context:
    """One HIP device + stream + workspace.  Not thread-safe: one per thread."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(_lib.cent_ctx_create(device, C.byref(h)), "cent_ctx_create",
               "needs a gfx950 (MI355X) device")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.cent_ctx_destroy(h)
            self._h = None

    def last_error(self) -> str:
        return _lib.cent_ctx_last_error(self._h).decode()

    def probe_copy_bandwidth(self, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """Read+write GB/s of a plain device copy (HBM ceiling / counter calibration)."""
        out = C.c_double()
        _check(_lib.cent_probe_copy_bandwidth(self._h, nbytes, iters, C.byref(out)),
               "cent_probe_copy_bandwidth", self.last_error())
        return out.value

    def probe_gather_bandwidth(self, piece: int, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """Useful GB/s of random gathers of ``piece``-byte aligned pieces (32 ... 512)."""
        out = C.c_double()
        _check(_lib.cent_probe_gather_bandwidth(self._h, nbytes, piece, iters, C.byref(out)),
               "cent_probe_gather_bandwidth", self.last_error())
        return out.value

    def probe_gather_lines(self, lines: int, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """Line GB/s of loads that each read ``lines`` (1, 2, 4) random whole 128-byte lines."""
        out = C.c_double()
        _check(_lib.cent_probe_gather_lines(self._h, nbytes, lines, iters, C.byref(out)),
               "cent_probe_gather_lines", self.last_error())
        return out.value

    def upload(self, csr: CSR, keep_order: bool = False, tuned: bool = True, order: str = "auto") -> DeviceGraph:
        return DeviceGraph(self, csr, keep_order, tuned, order)

    def tune(self, g: DeviceGraph, sources, what: int = CENT_TUNE_BC, max_trials: int = 0) -> TuneParams:
        """Autotune g's launch settings on the sample sources (at most 256 for
        betweenness, 1024 for closeness); applies and records the fastest
        (process-wide table) and returns it."""
        s = np.ascontiguousarray(sources, dtype=np.int32)
        p = TuneParams()
        _check(_lib.cent_tune(self._h, g._h, _ptr(s, C.c_int32), len(s), what, max_trials, C.byref(p)),
               "cent_tune", self.last_error())
        return p

    def workspace_bytes(self) -> int:
        """Device workspace the context holds (its second-stream peer included)."""
        return int(_lib.cent_ctx_workspace_bytes(self._h))

    def trim(self) -> None:
        """Frees the context's device workspace (it grows again on demand)."""
        _check(_lib.cent_ctx_trim(self._h), "cent_ctx_trim", self.last_error())

    def stats(self) -> dict:
        s = Stats()
        _check(_lib.cent_ctx_stats(self._h, C.byref(s)), "cent_ctx_stats")
        return s.as_dict()

    def reset_stats(self) -> None:
        _check(_lib.cent_ctx_reset_stats(self._h), "cent_ctx_reset_stats")

    @staticmethod
    def _src(g: DeviceGraph, sources):
        if sources is None:
            return None, -1, g.n
        s = np.ascontiguousarray(sources, dtype=np.int32)
        return s, len(s), len(s)

    def bc(self, g: DeviceGraph, sources=None, mode: int = CENT_BC_TEXTBOOK, bc=None,
           farness: bool = True, dist_sum: bool = False, timing: bool = False,
           grid_waves: int = 0, frontier: str = "auto", max_slots: int = 0,
           keep_source_order: bool = False, one_stream: bool = False, push_pct: int = 0):
        """Brandes BC from ``sources`` (None: all).  Returns (bc, farness, dist_sum).
        push_pct: level-2 push threshold of the dense step (0: tuned or default, -1: pull)."""
        s, k, kk = self._src(g, sources)
        out = np.zeros(g.n, np.float64) if bc is None else bc
        far = np.zeros(max(kk, 1), np.int64) if farness else None
        ds = np.zeros(max(g.n, 1), np.int64) if dist_sum else None
        o = _opts(mode, timing, grid_waves, frontier, max_slots, keep_source_order, one_stream, push_pct)
        _check(_lib.cent_bc(self._h, g._h, _ptr(s, C.c_int32) if s is not None and k else None, k,
                            _ptr(out, C.c_double) if g.n else None,
                            _ptr(far, C.c_int64) if far is not None else None,
                            _ptr(ds, C.c_int64) if ds is not None else None, C.byref(o)),
               "cent_bc", self.last_error())
        return out, (far[:kk] if far is not None else None), (ds[:g.n] if ds is not None else None)

    def cc(self, g: DeviceGraph, sources=None, dist_sum: bool = False, timing: bool = False,
           grid_waves: int = 0, frontier: str = "auto", max_slots: int = 0,
           keep_source_order: bool = False, one_stream: bool = False):
        """Closeness (farness) from ``sources`` (None: all).  Returns (farness, dist_sum)."""
        s, k, kk = self._src(g, sources)
        far = np.zeros(max(kk, 1), np.int64)
        ds = np.zeros(max(g.n, 1), np.int64) if dist_sum else None
        o = _opts(0, timing, grid_waves, frontier, max_slots, keep_source_order, one_stream)
        _check(_lib.cent_cc(self._h, g._h, _ptr(s, C.c_int32) if s is not None and k else None, k,
                            _ptr(far, C.c_int64),
                            _ptr(ds, C.c_int64) if ds is not None else None, C.byref(o)),
               "cent_cc", self.last_error())
        return far[:kk], (ds[:g.n] if ds is not None else None)

    def device_array(self, n: int, dtype, fill=None) -> "DeviceArray":
        """A device array of n elements on this context's GPU."""
        a = DeviceArray(self, n, dtype)
        if fill is not None:
            a.copy_from(np.full(n, fill, dtype=a.dtype))
        return a

    def _dev_arg(self, a, dtype, length: int, what: str):
        """Device address of a DeviceArray of this context with at least
        ``length`` elements of ``dtype`` (None passes through)."""
        if a is None:
            return None
        if not isinstance(a, DeviceArray) or a.ctx is not self:
            raise ValueError(f"{what}: need a DeviceArray of this context, got {type(a).__name__}")
        if a.dtype != np.dtype(dtype):
            raise ValueError(f"{what}: need {np.dtype(dtype)}, got {a.dtype}")
        if a.n < length:
            raise ValueError(f"{what}: {a.n} elements, need {length}")
        return a.ptr

    def bc_device(self, g: DeviceGraph, sources, bc, farness=None, dist_sum=None,
                  mode: int = CENT_BC_TEXTBOOK, timing: bool = False, grid_waves: int = 0,
                  frontier: str = "auto", max_slots: int = 0, keep_source_order: bool = False,
                  one_stream: bool = False) -> None:
        """Brandes BC with the results left in DeviceArrays of this context:
        bc (float64, n) accumulated, farness (int64, k) assigned, dist_sum
        (int64, n) accumulated; same meaning as bc()."""
        s, k, kk = self._src(g, sources)
        o = _opts(mode, timing, grid_waves, frontier, max_slots, keep_source_order, one_stream)
        _check(_lib.cent_bc_device(self._h, g._h, _ptr(s, C.c_int32) if s is not None and k else None, k,
                                   self._dev_arg(bc, np.float64, g.n, "bc"),
                                   self._dev_arg(farness, np.int64, kk, "farness"),
                                   self._dev_arg(dist_sum, np.int64, g.n, "dist_sum"), C.byref(o)),
               "cent_bc_device", self.last_error())

    def cc_device(self, g: DeviceGraph, sources, farness, dist_sum=None, timing: bool = False,
                  grid_waves: int = 0, frontier: str = "auto", max_slots: int = 0,
                  keep_source_order: bool = False, one_stream: bool = False) -> None:
        """Closeness with farness (int64, k) and dist_sum (int64, n) in
        DeviceArrays of this context; same meaning as cc()."""
        s, k, kk = self._src(g, sources)
        o = _opts(0, timing, grid_waves, frontier, max_slots, keep_source_order, one_stream)
        _check(_lib.cent_cc_device(self._h, g._h, _ptr(s, C.c_int32) if s is not None and k else None, k,
                                   self._dev_arg(farness, np.int64, kk, "farness"),
                                   self._dev_arg(dist_sum, np.int64, g.n, "dist_sum"), C.byref(o)),
               "cent_cc_device", self.last_error())

    def cc_device_noalloc(self, g: DeviceGraph, sources, farness=None, dist_sum=None) -> None:
        """cent_cc_device_noalloc: cc_device with every device buffer of the
        call in the library's fixed per-GPU scratch (nothing allocated)."""
        s, k, kk = self._src(g, sources)
        _check(_lib.cent_cc_device_noalloc(self._h, g._h, _ptr(s, C.c_int32) if s is not None and k else None, k,
                                           self._dev_arg(farness, np.int64, kk, "farness"),
                                           self._dev_arg(dist_sum, np.int64, g.n, "dist_sum")),
               "cent_cc_device_noalloc", self.last_error())

    def cc_reduced(self, red: Reduced, sched: int = CENT_SCHED_PERSISTENT, timing: bool = False):
        """Farness of every vertex through the D1/AP-reduced form."""
        out = np.zeros(max(red.n, 1), np.int64)
        o = _opts(0, timing)
        _check(_lib.cent_cc_reduced(self._h, red._h, _ptr(out, C.c_int64), sched, C.byref(o)),
               "cent_cc_reduced", self.last_error())
        return out[:red.n]

    def bc_reduced_sources(self, red: Reduced, sources, sched: int = CENT_SCHED_PERSISTENT, timing: bool = False):
        """(bc, farness of each source): betweenness from the source multiset
        through the reduced form (cent_bc_reduced_sources)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        bc = np.zeros(max(red.n, 1))
        far = np.zeros(max(len(src), 1), np.int64)
        o = _opts(0, timing)
        _check(_lib.cent_bc_reduced_sources(self._h, red._h, _ptr(src, C.c_int32) if len(src) else None, len(src),
                                            _ptr(bc, C.c_double), _ptr(far, C.c_int64), sched, C.byref(o)),
               "cent_bc_reduced_sources", self.last_error())
        return bc[:red.n], far[:len(src)]

    def bc_reduced(self, red: Reduced, sched: int = CENT_SCHED_PERSISTENT, timing: bool = False,
                   push_pct: int = 0):
        """(bc, farness) of every vertex, all sources, through the reduced form."""
        bc = np.zeros(max(red.n, 1))
        far = np.zeros(max(red.n, 1), np.int64)
        o = _opts(0, timing, push_pct=push_pct)
        _check(_lib.cent_bc_reduced(self._h, red._h, _ptr(bc, C.c_double), _ptr(far, C.c_int64),
                                    sched, C.byref(o)), "cent_bc_reduced", self.last_error())
        return bc[:red.n], far[:red.n]


_lib.cent_debug_fail_upload.argtypes = [C.c_int32]


def debug_fail_upload(nth: int) -> None:
    """Test hook: the nth device upload of a graph structure from now on fails
    with CENT_E_OOM (0 switches the injection off)."""
    _check(_lib.cent_debug_fail_upload(nth), "cent_debug_fail_upload")




# ---------------------------------------------------------------- serving
CENT_SERVER_DRYRUN = 1
CENT_Q_BC, CENT_Q_CC = 0, 1
REQ_STATE = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


#: Source-parallel requests run by the server use at most this many per-source
#: slots (bc_amd/csrc/server.cpp kSourceSlots; the admission estimate counts them).
SERVER_SOURCE_SLOTS = 256


class _ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("hbm_budget", C.c_int64),
                ("max_queue", C.c_int32), ("flags", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("status", C.c_int32), ("slot", C.c_int32),
                ("reserved", C.c_int32), ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_end_ms", C.c_double), ("footprint", C.c_int64)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}
        d["state"] = REQ_STATE.get(self.state, self.state)
        return d


class ServerStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in (
        "submitted", "completed", "failed", "cancelled", "rejected", "queued", "running",
        "max_queued", "max_running", "graph_uploads", "graph_hits", "graph_evictions",
        "used_workspace", "used_graph")]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib.cent_server_create.argtypes = [C.POINTER(_ServerOpts), C.POINTER(C.c_void_p)]
_lib.cent_server_destroy.argtypes = [C.c_void_p]
_lib.cent_server_destroy.restype = None
_lib.cent_server_add_graph.argtypes = [C.c_void_p, _P_CSR, C.POINTER(C.c_int32)]
_lib.cent_server_submit.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _I32P, C.c_int32,
                                    C.POINTER(C.c_int64)]
_lib.cent_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(RequestInfo)]
_lib.cent_server_result.argtypes = [C.c_void_p, C.c_int64, _F64P, _I64P, C.POINTER(RequestInfo)]
_lib.cent_server_cancel.argtypes = [C.c_void_p, C.c_int64]
_lib.cent_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]


class Server:
    """Continuous-batching query server (see cent_server_* in include/cent.h)."""

    def __init__(self, slots: int = 4, hbm_budget: int = 64 << 30, max_queue: int = 100000,
                 device: int = 0, dryrun: bool = False):
        o = _ServerOpts(device, slots, hbm_budget, max_queue, CENT_SERVER_DRYRUN if dryrun else 0)
        h = C.c_void_p()
        _check(_lib.cent_server_create(C.byref(o), C.byref(h)), "cent_server_create")
        self._h = h
        self._graphs = []
        self._meta = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.cent_server_destroy(h)
            self._h = None

    __del__ = close

    def add_graph(self, csr: CSR) -> int:
        gid = C.c_int32()
        _check(_lib.cent_server_add_graph(self._h, csr._p, C.byref(gid)), "cent_server_add_graph")
        self._graphs.append(csr.n)
        return gid.value

    def submit(self, gid: int, kind: int, sources=None) -> int:
        rid = C.c_int64()
        if sources is None:
            st = _lib.cent_server_submit(self._h, gid, kind, None, -1, C.byref(rid))
            k = -1
        else:
            src = np.ascontiguousarray(sources, dtype=np.int32)
            st = _lib.cent_server_submit(self._h, gid, kind, _ptr(src, C.c_int32) if len(src) else None,
                                         len(src), C.byref(rid))
            k = len(src)
        _check(st, "cent_server_submit")
        self._meta[rid.value] = (gid, kind, k)
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> dict:
        info = RequestInfo()
        _check(_lib.cent_server_wait(self._h, rid, timeout_ms, C.byref(info)), "cent_server_wait")
        return info.as_dict()

    def result(self, rid: int):
        """(status, bc or None, farness, info) — forgets the request."""
        gid, kind, k = self._meta.pop(rid)
        n = self._graphs[gid]
        bc = np.zeros(max(n, 1)) if kind == CENT_Q_BC else None
        far = np.zeros(max(n if k == -1 else k, 1), np.int64)
        info = RequestInfo()
        st = _lib.cent_server_result(self._h, rid, _ptr(bc, C.c_double) if bc is not None else None,
                                     _ptr(far, C.c_int64), C.byref(info))
        return st, (bc[:n] if bc is not None else None), far[:(n if k == -1 else k)], info.as_dict()

    def cancel(self, rid: int) -> None:
        _check(_lib.cent_server_cancel(self._h, rid), "cent_server_cancel")

    def stats(self) -> dict:
        s = ServerStats()
        _check(_lib.cent_server_get_stats(self._h, C.byref(s)), "cent_server_get_stats")
        return s.as_dict()


# ------------------------------------------- reference-shaped entry points
def compute_D1_AP_CC_shareBased_DegreeOrder(csr: CSR, CCs: np.ndarray) -> np.ndarray:
    """Farness of every vertex via degree-1 folding + articulation-point split,
    assigned into CCs (int64) like Seq_CC/CC.cpp:1728-2052."""
    if CCs.dtype != np.int64 or len(CCs) < csr.n or not CCs.flags.c_contiguous:
        raise ValueError("CCs must be a contiguous int64 array of length n")
    _check(_lib.compute_D1_AP_CC_shareBased_DegreeOrder(csr._p, _ptr(CCs, C.c_int64)),
           "compute_D1_AP_CC_shareBased_DegreeOrder")
    return CCs


def computeBC(csr: CSR, BCs: np.ndarray, compat: bool = False) -> np.ndarray:
    """Brandes BC over sources startNodeID..endNodeID, accumulated into BCs
    (float64, length n) like Seq_Brandes/Brandes.c:224 accumulates into its
    caller-zeroed float array.  compat=True uses the reference's per-predecessor
    accumulation (Brandes.c:317-327)."""
    if BCs.dtype != np.float64 or len(BCs) < csr.n or not BCs.flags.c_contiguous:
        raise ValueError("BCs must be a contiguous float64 array of length n")
    fn = _lib.computeBC_compat if compat else _lib.computeBC
    _check(fn(csr._p, _ptr(BCs, C.c_double)), "computeBC")
    return BCs


def computeCC(csr: CSR, CCs: np.ndarray) -> np.ndarray:
    """Farness of every source startNodeID..endNodeID, assigned into CCs[id]
    (int64), like Seq_CC/CC.cpp:37-117."""
    if CCs.dtype != np.int64 or len(CCs) < csr.n or not CCs.flags.c_contiguous:
        raise ValueError("CCs must be a contiguous int64 array of length n")
    _check(_lib.computeCC(csr._p, _ptr(CCs, C.c_int64)), "computeCC")
    return CCs


__all__ = [
    "CSR", "CentError", "Context", "DeviceGraph", "buildGraph", "createCSR", "computeBC",
    "computeCC", "compute_D1_AP_CC_shareBased_DegreeOrder", "Reduced", "Server", "gen_rmat", "gen_grid", "gen_blocktree", "sample_sources", "LIB_PATH",
    "CENT_BC_TEXTBOOK", "CENT_BC_REF_COMPAT",
]
