"""ctypes bindings of the CPU oracle (test infrastructure only).

The oracle is the checker for the HIP library: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The product package (cytom_amd) never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

BFS, CC, SSSP, PRD, PPR = 0, 1, 2, 3, 4
KIND_NAMES = {"bfs": BFS, "cc": CC, "sssp": SSSP, "pagerank-delta": PRD, "prd": PRD, "ppr": PPR}

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("converged", ctypes.c_uint32),
        ("reexecuted", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("traversed_edges", ctypes.c_uint64),
        ("active_vertices", ctypes.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("oracle library not built: run `make -C oracle`")
    L = ctypes.CDLL(LIB_PATH)
    L.orc_hash32.argtypes = [ctypes.c_uint32]
    L.orc_hash32.restype = ctypes.c_uint32
    L.orc_hash64.argtypes = [ctypes.c_uint64]
    L.orc_hash64.restype = ctypes.c_uint64
    L.orc_hash_double.argtypes = [ctypes.c_uint64]
    L.orc_hash_double.restype = ctypes.c_double
    for name in ("orc_rmat_edges", "orc_rmat_edges_recursive"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                      _u32p, _u32p]
        f.restype = None
    L.orc_rmat_gather.argtypes = [ctypes.c_uint32, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                  _u32p, _u32p]
    L.orc_rmat_gather.restype = None
    L.orc_stream_weights_u32_at.argtypes = [ctypes.c_uint64, _u64p, ctypes.c_uint64, _u32p]
    L.orc_stream_weights_u32_at.restype = None
    L.orc_stream_weights_u32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _u32p]
    L.orc_stream_weights_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _f32p]
    L.orc_stream_create.argtypes = [ctypes.c_uint64]
    L.orc_stream_create.restype = ctypes.c_void_p
    L.orc_stream_destroy.argtypes = [ctypes.c_void_p]
    L.orc_stream_next.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.POINTER(ctypes.c_uint64), _u64p]
    L.orc_stream_next.restype = ctypes.c_int64
    L.orc_stream_live_count.argtypes = [ctypes.c_void_p]
    L.orc_stream_live_count.restype = ctypes.c_uint64
    L.orc_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                             ctypes.c_double, ctypes.c_int]
    L.orc_create.restype = ctypes.c_void_p
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_num_vertices.argtypes = [ctypes.c_void_p]
    L.orc_num_vertices.restype = ctypes.c_uint64
    L.orc_register.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
    L.orc_register.restype = ctypes.c_int
    L.orc_update.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_uint64]
    L.orc_update.restype = ctypes.c_int64
    L.orc_execute.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Stats)]
    L.orc_execute.restype = ctypes.c_int
    L.orc_read.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.orc_read.restype = ctypes.c_int
    L.orc_last_reexec.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_last_reexec.restype = ctypes.c_int
    L.orc_live_edges.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    L.orc_live_edges.restype = ctypes.c_uint64
    L.orc_degrees.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    L.orc_check_bfs.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, ctypes.c_uint32,
                                _u32p]
    L.orc_check_sssp.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, _u32p,
                                 ctypes.c_uint32, _u32p]
    L.orc_check_cc.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, _u32p]
    L.orc_check_multi.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p,
                                  ctypes.c_void_p, _u32p, ctypes.c_uint32, ctypes.c_int, _u32p]
    L.orc_check_multi.restype = None
    L.orc_checker_threads.argtypes = [ctypes.c_int]
    L.orc_checker_threads.restype = None
    L.orc_check_cc_mt.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, ctypes.c_int,
                                  _u32p]
    L.orc_prd_f64_mt.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p,
                                 ctypes.c_uint32, ctypes.c_double, ctypes.c_int, _f64p]
    L.orc_prd_f64_mt.restype = ctypes.c_uint32
    L.orc_prd_residual.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, _f32p,
                                   ctypes.c_double, ctypes.c_int, _f64p]
    L.orc_prd_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, ctypes.c_uint32,
                              ctypes.c_double, _f64p]
    L.orc_prd_f64.restype = ctypes.c_uint32
    L.orc_set_sources.argtypes = [ctypes.c_void_p, ctypes.c_int, _u32p, ctypes.c_uint32]
    L.orc_set_sources.restype = ctypes.c_int
    L.orc_ppr_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p, _u32p,
                              ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                              _f64p]
    L.orc_ppr_f64.restype = ctypes.c_uint32
    L.orc_pagerank_f64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_uint32, _f64p]
    L.orc_pagerank_f64.restype = ctypes.c_uint32
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def hash64(x):
    return lib().orc_hash64(x)


def hash_double(x):
    return lib().orc_hash_double(x)


def rmat_edges(scale, seed, first, n, recursive=False):
    s = np.empty(n, np.uint32)
    d = np.empty(n, np.uint32)
    f = lib().orc_rmat_edges_recursive if recursive else lib().orc_rmat_edges
    f(scale, seed, first, n, s, d)
    return s, d


def weights_u32(seed, first, n):
    w = np.empty(n, np.uint32)
    lib().orc_stream_weights_u32(seed, first, n, w)
    return w


def weights_f32(seed, first, n):
    w = np.empty(n, np.float32)
    lib().orc_stream_weights_f32(seed, first, n, w)
    return w


class Stream:
    """Update stream: consecutive RMAT inserts, deletions sampled from live."""

    def __init__(self, seed):
        self._h = lib().orc_stream_create(seed)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_stream_destroy(self._h)
            self._h = None

    def next(self, n_ins, n_del):
        first = ctypes.c_uint64(0)
        idx = np.empty(max(n_del, 1), np.uint64)
        m = lib().orc_stream_next(self._h, n_ins, n_del, ctypes.byref(first), idx)
        return first.value, idx[:m].copy()

    @property
    def live(self):
        return lib().orc_stream_live_count(self._h)


def batch_edges(scale, seed, first, n_ins, del_idx, weighted):
    """Materialise one batch's (src, dst, w) arrays from stream indices."""
    s, d = rmat_edges(scale, seed, first, n_ins)
    w = weights_u32(seed, first, n_ins) if weighted else None
    idx = np.ascontiguousarray(del_idx, dtype=np.uint64)
    ds = np.empty(len(idx), np.uint32)
    dd = np.empty(len(idx), np.uint32)
    lib().orc_rmat_gather(scale, seed, idx, len(idx), ds, dd)
    dw = None
    if weighted:
        dw = np.empty(len(idx), np.uint32)
        lib().orc_stream_weights_u32_at(seed, idx, len(idx), dw)
    return s, d, w, ds, dd, dw


class Graph:
    """Oracle tile store + incremental executors."""

    def __init__(self, n_vertices, vertices_per_tile=65536, weighted=False,
                 max_iterations=100, delta=1e-7, threads=1):
        self._h = lib().orc_create(n_vertices, vertices_per_tile, int(weighted), max_iterations,
                                   delta, threads)
        if not self._h:
            raise ValueError("invalid oracle configuration")
        self.V = lib().orc_num_vertices(self._h)
        self.weighted = weighted
        self.kinds = []

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def register(self, kind, source=0):
        pid = lib().orc_register(self._h, kind, source)
        if pid < 0:
            raise ValueError("register failed: %d" % pid)
        self.kinds.append(kind)
        return pid

    def update(self, src, dst, w=None, dsrc=None, ddst=None, dw=None):
        src, dst, w = _u32(src), _u32(dst), _u32(w)
        dsrc = _u32(dsrc if dsrc is not None else np.empty(0, np.uint32))
        ddst = _u32(ddst if ddst is not None else np.empty(0, np.uint32))
        dw = _u32(dw)
        r = lib().orc_update(self._h, _ptr(src), _ptr(dst), _ptr(w), len(src), _ptr(dsrc),
                             _ptr(ddst), _ptr(dw), len(dsrc))
        if r < 0:
            raise ValueError("update failed: %d" % r)
        return r

    def execute(self, pid):
        st = Stats()
        r = lib().orc_execute(self._h, pid, ctypes.byref(st))
        if r < 0:
            raise ValueError("execute failed")
        return st

    def set_sources(self, pid, sources):
        s = _u32(np.asarray(sources, np.uint32))
        if lib().orc_set_sources(self._h, pid, s, len(s)) < 0:
            raise ValueError("set_sources failed")

    def read(self, pid):
        dt = np.float32 if self.kinds[pid] in (PRD, PPR) else np.uint32
        out = np.empty(self.V, dt)
        lib().orc_read(self._h, pid, _ptr(out))
        return out

    def last_reexec(self, pid):
        return lib().orc_last_reexec(self._h, pid)

    def live_edges(self):
        n = lib().orc_live_edges(self._h, None, None, None)
        s = np.empty(n, np.uint32)
        d = np.empty(n, np.uint32)
        w = np.empty(n, np.uint32)
        lib().orc_live_edges(self._h, _ptr(s), _ptr(d), _ptr(w))
        return s, d, w

    def degrees(self):
        o = np.empty(self.V, np.uint32)
        i = np.empty(self.V, np.uint32)
        lib().orc_degrees(self._h, o, i)
        return o, i


def checker_threads(n):
    """Threads the textbook checkers use to build their rows."""
    lib().orc_checker_threads(int(n))


def check_bfs(V, s, d, source):
    out = np.empty(V, np.uint32)
    lib().orc_check_bfs(V, len(s), _u32(s), _u32(d), source, out)
    return out


def check_sssp(V, s, d, w, source):
    out = np.empty(V, np.uint32)
    lib().orc_check_sssp(V, len(s), _u32(s), _u32(d), _u32(w), source, out)
    return out


def check_multi(V, s, d, sources, w=None, threads=1):
    """Textbook BFS levels (w None) or Dijkstra distances from each source:
    an (n, V) uint32 array."""
    src = _u32(np.atleast_1d(sources))
    out = np.empty((len(src), V), np.uint32)
    w = _u32(w)
    lib().orc_check_multi(V, len(s), _u32(s), _u32(d), _ptr(w), src, len(src), threads, out)
    return out


def check_cc(V, s, d, threads=1):
    out = np.empty(V, np.uint32)
    if threads > 1:
        lib().orc_check_cc_mt(V, len(s), _u32(s), _u32(d), threads, out)
    else:
        lib().orc_check_cc(V, len(s), _u32(s), _u32(d), out)
    return out


def prd_f64(V, s, d, max_iterations=100, delta=1e-7, threads=1):
    out = np.empty(V, np.float64)
    it = lib().orc_prd_f64_mt(V, len(s), _u32(s), _u32(d), max_iterations, delta, threads, out)
    return out, it


def prd_residual(V, s, d, rank, alpha=0.85, threads=1):
    """Normalised one-sweep residual of the PageRank fixed point over FP32
    ranks: sum|r - (1-a) - a P^T r| / sum|r| (FP64)."""
    out = np.zeros(2, np.float64)
    r = np.ascontiguousarray(rank, dtype=np.float32)
    lib().orc_prd_residual(V, len(s), _u32(s), _u32(d), r, alpha, threads, out)
    return out[0] / out[1] if out[1] else 0.0


def host_threads():
    """Host threads a checker may use: the CPUs this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it."""
    import os
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def ppr_f64(V, s, d, teleport, alpha=0.85, tol=1e-13, max_iterations=10000):
    t = _u32(np.asarray(teleport, np.uint32))
    out = np.empty(V, np.float64)
    it = lib().orc_ppr_f64(V, len(s), _u32(s), _u32(d), t, len(t), alpha, tol, max_iterations,
                           out)
    return out, it


def pagerank_f64(V, s, d, alpha=0.85, tol=1e-12, max_iterations=10000):
    out = np.empty(V, np.float64)
    it = lib().orc_pagerank_f64(V, len(s), _u32(s), _u32(d), alpha, tol, max_iterations, out)
    return out, it
