"""Python host side of the MI355X-native evolving-graph engine.

Thin ctypes layer over the C-ABI in include/cytom.h (libcytom_hip.so, built
in-tree by cytom_amd/build.py).  The vertex programs keep the reference's
algorithm-store names (examples/bfs.cc main_bfs, cc.cc main_cc, sssp.cc
main_sssp, pagerank-delta.cc main_pagerank_delta).  There is no CPU
fallback: if the HIP library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

from .formats import PRD_VERTEX as PRD_RECORD  # results-file vertex of PRD/PPR

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcytom_hip.so")

BFS, CC, SSSP, PRD, PPR = 0, 1, 2, 3, 4
PROGRAMS = {"bfs": BFS, "cc": CC, "sssp": SSSP, "pagerank-delta": PRD, "ppr": PPR}
RESORT_OFF = 0xFFFFFFFF  # cytom_tuning.resort_pct value that disables sort-compaction

EINVAL, ENOMEM, ERANGE, EDEVICE = -22, -12, -34, -5


class CytomError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cytom error %d: %s" % (code, msg))
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("n_vertices", ctypes.c_uint64),
        ("vertices_per_tile", ctypes.c_uint32),
        ("max_iterations", ctypes.c_uint32),
        ("delta", ctypes.c_double),
        ("edge_capacity", ctypes.c_uint64),
        ("weighted", ctypes.c_uint32),
        ("device", ctypes.c_uint32),
    ]


STATICS_GIVEN = 1 << 0
NEEDS_SOURCES_ACTIVE = 1 << 1
NEEDS_TARGETS_ACTIVE = 1 << 2
ASYNC = 1 << 3
USE_CHANGED = 1 << 4
FIRST_ITERATION_NONE, FIRST_ITERATION_PRD_ALL = 0, 1


class ProgramDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("source", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("vertex_bytes", ctypes.c_uint32),
                ("first_iteration", ctypes.c_uint32), ("n_fields", ctypes.c_uint32),
                ("neutral_element", ctypes.c_double * 3)]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("converged", ctypes.c_uint32),
        ("reexecuted", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("traversed_edges", ctypes.c_uint64),
        ("active_vertices", ctypes.c_uint64),
        ("slots_scanned", ctypes.c_uint64),
        ("t_compute_ms", ctypes.c_double),
    ]


class Query(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("n_sources", ctypes.c_uint32),
        ("sources", ctypes.c_void_p),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("chunk_slots", ctypes.c_uint32),
        ("job_chunks", ctypes.c_uint32),
        ("blocks_per_cu", ctypes.c_uint32),
        ("source_bands", ctypes.c_uint32),
        ("resort_pct", ctypes.c_uint32),
        ("resort_max_slots", ctypes.c_uint32),
        ("prd_frac_bits", ctypes.c_uint32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("inserted", ctypes.c_uint64),
        ("deleted", ctypes.c_uint64),
        ("missing", ctypes.c_uint64),
        ("live_edges", ctypes.c_uint64),
        ("tiles_relocated", ctypes.c_uint32),
        ("tiles_compacted", ctypes.c_uint32),
        ("tiles_sorted", ctypes.c_uint32),
        ("host_syncs", ctypes.c_uint32),
        ("t_ingest_ms", ctypes.c_double),
    ]


# Every function include/cytom.h declares: name -> (restype, argtypes).
_VP = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
SIGNATURES = {
    "cytom_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_VP)]),
    "cytom_destroy": (ctypes.c_int, [_VP]),
    "cytom_last_error": (ctypes.c_char_p, [_VP]),
    "cytom_geometry": (ctypes.c_int, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U32),
                                      ctypes.POINTER(_U32)]),
    "cytom_program_register": (ctypes.c_int, [_VP, ctypes.POINTER(ProgramDesc),
                                              ctypes.POINTER(ctypes.c_int)]),
    "cytom_program_init": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cytom_program_statics": (ctypes.c_int, [_U32, ctypes.POINTER(ProgramDesc)]),
    "cytom_program_set_sources": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _U32]),
    "cytom_batched_execute": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "cytom_ppr_weighted": (ctypes.c_int, [_U64, _U64, _VP, _VP, _VP, _VP, ctypes.c_float,
                                          ctypes.c_float, _U32, _VP, _VP]),
    "cytom_ppr_weighted_noalloc": (ctypes.c_int, [_U64, _U64, _VP, _VP, _VP, _VP,
                                                  ctypes.c_float, ctypes.c_float, _U32, _VP,
                                                  _VP]),
    "cytom_ppr_adjoint": (ctypes.c_int, [_U64, _U64, _VP, _VP, _VP, _VP, _VP, ctypes.c_float,
                                         ctypes.c_float, _U32, _VP, _VP, _VP]),
    "cytom_ppr_graph_create": (ctypes.c_int, [_U64, _U64, _VP, _VP, _VP,
                                              ctypes.POINTER(ctypes.c_void_p)]),
    "cytom_ppr_graph_set_weights": (ctypes.c_int, [_VP, _VP]),
    "cytom_ppr_graph_forward": (ctypes.c_int, [_VP, _VP, ctypes.c_float, ctypes.c_float, _U32,
                                               _VP, _VP]),
    "cytom_ppr_graph_adjoint": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_float, ctypes.c_float,
                                               _U32, _VP, _VP, _VP]),
    "cytom_ppr_graph_destroy": (ctypes.c_int, [_VP]),
    "cytom_ppr_last_error": (ctypes.c_char_p, []),
    "cytom_read_query": (ctypes.c_int, [_VP, _U32, _VP, ctypes.c_size_t]),
    "cytom_update_batch": (ctypes.c_int, [_VP, _VP, _VP, _VP, _U64, _VP, _VP, _VP, _U64,
                                          ctypes.POINTER(BatchStats)]),
    "cytom_update_batch_device": (ctypes.c_int, [_VP, _VP, _VP, _VP, _U64, _VP, _VP, _VP, _U64,
                                                 ctypes.POINTER(BatchStats)]),
    "cytom_update_batch_records": (ctypes.c_int, [_VP, _VP, _U32, _U64, _U32, _VP, _U64,
                                                  ctypes.POINTER(BatchStats)]),
    "cytom_read_results": (ctypes.c_int64, [_VP, ctypes.c_int, _VP, ctypes.c_size_t]),
    "cytom_execute": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(RunStats)]),
    "cytom_execute_many": (ctypes.c_int, [_VP, _VP, _U32, _VP]),
    "cytom_ppr_program_adjoint": (ctypes.c_int64, [_VP, ctypes.c_int, _VP, ctypes.c_float, _U32,
                                                 _VP, _VP, _VP, _VP, _U64,
                                                 ctypes.POINTER(ctypes.c_uint32)]),
    "cytom_step": (ctypes.c_int, [_VP, _VP, _VP, _VP, _U64, _VP, _VP, _VP, _U64, _VP, _U32,
                                  ctypes.POINTER(BatchStats), _VP]),
    "cytom_read_values": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.c_size_t]),
    "cytom_read_degrees": (ctypes.c_int, [_VP, _VP, _VP]),
    "cytom_export_edges": (ctypes.c_int64, [_VP, _VP, _VP, _VP, _U64]),
    "cytom_set_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    "cytom_get_tuning": (ctypes.c_int, [_VP, ctypes.POINTER(Tuning)]),
    "cytom_set_profiling": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cytom_set_step_graphs": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cytom_set_concurrent_runs": (ctypes.c_int, [_VP, ctypes.c_int]),
    "cytom_step_graph_stats": (ctypes.c_int, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U64),
                                              ctypes.POINTER(ctypes.c_uint32)]),
    "cytom_kernel_profile": (ctypes.c_int, [_VP, _U32, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(_U64),
                                            ctypes.POINTER(ctypes.c_double)]),
    "cytom_rmat_generate_device": (ctypes.c_int, [_U32, _U64, _U64, _U64, _VP, _VP, _VP]),
    "cytom_rmat_gather_device": (ctypes.c_int, [_U32, _U64, _VP, _U64, _VP, _VP, _VP]),
    "cytom_stream_create": (ctypes.c_int, [_U64, ctypes.POINTER(_VP)]),
    "cytom_stream_destroy": (ctypes.c_int, [_VP]),
    "cytom_stream_next": (ctypes.c_int64, [_VP, _U64, _U64, ctypes.POINTER(_U64), _VP]),
    "cytom_device_alloc": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_size_t]),
    "cytom_device_free": (ctypes.c_int, [_VP]),
    "cytom_memcpy_h2d": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
    "cytom_memcpy_d2h": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
    "cytom_memcpy_d2d": (ctypes.c_int, [_VP, _VP, ctypes.c_size_t]),
    "cytom_device_synchronize": (ctypes.c_int, []),
    "cytom_device_count": (ctypes.c_int, []),
}

_lib = None


def lib():
    """Load libcytom_hip.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcytom_hip.so is not built: run `python -m cytom_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data_as(ctypes.c_void_p)


def _u32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.uint32)


def _ids(a, name):
    """Vertex ids as a contiguous 1-D uint32 array; rejects values that do not
    fit in 32 bits (negative or above 2^32 - 1) instead of wrapping them."""
    a = np.asarray(a)
    if a.ndim != 1:
        raise ValueError("%s must be one-dimensional" % name)
    if a.size and a.dtype.kind in "iuf":
        if a.dtype.kind == "f" and not np.array_equal(a, np.floor(a)):
            raise ValueError("%s holds non-integral ids" % name)
        if a.min() < 0 or a.max() > 0xFFFFFFFF:
            raise ValueError("%s holds ids outside [0, 2^32)" % name)
    elif a.size:
        raise ValueError("%s must be numeric" % name)
    return np.ascontiguousarray(a, dtype=np.uint32)


def _weights(a, name):
    """Edge weights as uint32; they must be integral and in [1, 2^32) (the
    engine's u32 SSSP weights), so a float weight is never truncated."""
    if a is None:
        return None
    a = _ids(a, name)
    if a.size and a.min() == 0:
        raise ValueError("%s holds zero weights (weights must be >= 1)" % name)
    return a


def _same_len(*named):
    arrs = [(n, a) for n, a in named if a is not None]
    if len(set(len(a) for _, a in arrs)) > 1:
        raise ValueError("array lengths differ: " +
                         ", ".join("%s=%d" % (n, len(a)) for n, a in arrs))




def _records(a):
    """Host edge records as a C-contiguous array of 8- or 12-byte items."""
    a = np.asarray(a)
    if a.dtype.names is None:
        if a.ndim != 2 or a.shape[1] not in (2, 3) or a.dtype.itemsize != 4:
            raise ValueError("records must be (n, 2|3) 4-byte arrays or a record dtype")
        a = np.ascontiguousarray(a).view(np.dtype((np.void, 4 * a.shape[1]))).reshape(-1)
    elif a.dtype.itemsize not in (8, 12):
        raise ValueError("record size must be 8 or 12 bytes")
    return np.ascontiguousarray(a)


def device_count():
    return lib().cytom_device_count()


def synchronize():
    rc = lib().cytom_device_synchronize()
    if rc:
        raise CytomError(rc, "device synchronize failed")


class DeviceArray:
    """A flat device allocation (HBM) with host copy helpers."""

    def __init__(self, n, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = lib().cytom_device_alloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc:
            raise CytomError(rc, "device allocation of %d bytes failed" % self.nbytes)
        self._ptr = p

    @property
    def ptr(self):
        return self._ptr

    def offset(self, elems):
        """Raw pointer to element `elems` (for sub-range views)."""
        return ctypes.c_void_p(self._ptr.value + int(elems) * self.dtype.itemsize)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            rc = lib().cytom_memcpy_h2d(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)
            if rc:
                raise CytomError(rc, "H2D copy failed")
        return d

    def copy_from(self, src, n=None):
        """Copies the first n elements (all of src by default) of the device
        array src into this one."""
        n = src.n if n is None else int(n)
        if n > self.n or n > src.n:
            raise ValueError("copy of %d elements between arrays of %d and %d" % (n, src.n, self.n))
        if n:
            rc = lib().cytom_memcpy_d2d(self.ptr, src.ptr, n * self.dtype.itemsize)
            if rc:
                raise CytomError(rc, "D2D copy failed")

    def to_host(self, n=None):
        n = self.n if n is None else int(n)
        out = np.empty(n, self.dtype)
        if n:
            rc = lib().cytom_memcpy_d2h(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                        n * self.dtype.itemsize)
            if rc:
                raise CytomError(rc, "D2H copy failed")
        return out

    def free(self):
        if getattr(self, "_ptr", None) is not None and self._ptr.value:
            lib().cytom_device_free(self._ptr)
            self._ptr = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rmat_generate(scale, seed, first, n, weights=False):
    """Device RMAT edges [first, first+n): returns DeviceArrays (src, dst, w|None)."""
    s = DeviceArray(n, np.uint32)
    d = DeviceArray(n, np.uint32)
    w = DeviceArray(n, np.uint32) if weights else None
    rc = lib().cytom_rmat_generate_device(scale, seed, first, n, s.ptr, d.ptr,
                                          w.ptr if w else None)
    if rc:
        raise CytomError(rc, "rmat generation failed")
    return s, d, w


def rmat_gather(scale, seed, index_dev, n, weights=False):
    s = DeviceArray(n, np.uint32)
    d = DeviceArray(n, np.uint32)
    w = DeviceArray(n, np.uint32) if weights else None
    rc = lib().cytom_rmat_gather_device(scale, seed, index_dev.ptr if n else None, n,
                                        s.ptr, d.ptr, w.ptr if w else None)
    if rc:
        raise CytomError(rc, "rmat gather failed")
    return s, d, w


class Stream:
    """Synthetic update stream: consecutive RMAT inserts, sampled deletions."""

    def __init__(self, seed):
        h = ctypes.c_void_p()
        rc = lib().cytom_stream_create(seed, ctypes.byref(h))
        if rc:
            raise CytomError(rc, "stream create failed")
        self._h = h

    def next(self, n_ins, n_del):
        first = ctypes.c_uint64(0)
        idx = np.empty(max(int(n_del), 1), np.uint64)
        m = lib().cytom_stream_next(self._h, n_ins, n_del, ctypes.byref(first),
                                    idx.ctypes.data_as(ctypes.c_void_p))
        if m < 0:
            raise CytomError(m, "stream next failed")
        return first.value, idx[:m].copy()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().cytom_stream_destroy(self._h)
                self._h = None
        except Exception:  # interpreter shutdown: the binding may already be gone
            pass


def program_statics(kind):
    """The Algorithm concept's statics of a kind, as a dict (the fields of
    cytom_program_desc after the kind and source)."""
    if isinstance(kind, str):
        kind = PROGRAMS[kind]
    d = ProgramDesc()
    rc = lib().cytom_program_statics(kind, ctypes.byref(d))
    if rc < 0:
        raise CytomError(rc, "program_statics: unknown kind %r" % (kind,))
    return {"flags": d.flags, "vertex_bytes": d.vertex_bytes,
            "first_iteration": d.first_iteration, "n_fields": d.n_fields,
            "neutral_element": list(d.neutral_element)}


class Graph:
    """One evolving graph in HBM plus the vertex programs registered on it."""

    def __init__(self, n_vertices, vertices_per_tile=0, weighted=False, max_iterations=100,
                 delta=1e-7, edge_capacity=0, device=0):
        cfg = Config(n_vertices, vertices_per_tile, max_iterations, delta, edge_capacity,
                     1 if weighted else 0, device)
        h = ctypes.c_void_p()
        rc = lib().cytom_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise CytomError(rc, "cytom_create failed")
        self._h = h
        V = ctypes.c_uint64()
        vpt = ctypes.c_uint32()
        T = ctypes.c_uint32()
        lib().cytom_geometry(h, ctypes.byref(V), ctypes.byref(vpt), ctypes.byref(T))
        self.V, self.vpt, self.T = V.value, vpt.value, T.value
        self.weighted = weighted
        self.kinds = []

    def _check(self, rc, what):
        if rc < 0:
            msg = lib().cytom_last_error(self._h)
            raise CytomError(rc, "%s: %s" % (what, msg.decode() if msg else ""))
        return rc

    def close(self):
        if getattr(self, "_h", None):
            lib().cytom_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register(self, kind, source=0, statics=None):
        """Registers a program of kind (name or number) with its source.
        statics: optional dict of the Algorithm concept's statics
        (program_statics(kind) gives the kind's own); registration refuses
        statics the kind's kernels do not implement."""
        if isinstance(kind, str):
            kind = PROGRAMS[kind]
        desc = ProgramDesc(kind, source)
        if statics is not None:
            for k, v in statics.items():
                if k == "neutral_element":
                    for i, x in enumerate(v):
                        desc.neutral_element[i] = x
                else:
                    setattr(desc, k, v)
            desc.flags |= STATICS_GIVEN
        pid = ctypes.c_int()
        self._check(lib().cytom_program_register(self._h, ctypes.byref(desc), ctypes.byref(pid)),
                    "register")
        self.kinds.append(kind)
        return pid.value

    def init(self, pid):
        self._check(lib().cytom_program_init(self._h, pid), "program init")

    def set_sources(self, pid, sources):
        """New BFS/SSSP source (one vertex) or PPR teleport set; re-inits."""
        s = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
        self._check(lib().cytom_program_set_sources(self._h, pid, s.ctypes.data, len(s)),
                    "set_sources")

    def update(self, src, dst, w=None, dsrc=None, ddst=None, dw=None):
        """One update batch from host arrays: inserts (src, dst[, w]) then
        deletions (dsrc, ddst[, dw]).  Arrays of one side must have equal
        lengths; weights must be positive integers below 2^32."""
        src, dst = _ids(src, "src"), _ids(dst, "dst")
        w = _weights(w, "w")
        dsrc = _ids(dsrc if dsrc is not None else np.empty(0, np.uint32), "dsrc")
        ddst = _ids(ddst if ddst is not None else np.empty(0, np.uint32), "ddst")
        dw = _weights(dw, "dw")
        _same_len(("src", src), ("dst", dst), ("w", w))
        _same_len(("dsrc", dsrc), ("ddst", ddst), ("dw", dw))
        st = BatchStats()
        self._check(lib().cytom_update_batch(self._h, _p(src), _p(dst), _p(w), len(src), _p(dsrc),
                                             _p(ddst), _p(dw), len(dsrc), ctypes.byref(st)),
                    "update")
        return st

    def update_device(self, src, dst, w, n_ins, dsrc=None, ddst=None, dw=None, n_del=0):
        """One update batch from device arrays (DeviceArray or raw pointers)
        of at least n_ins / n_del elements."""
        for name, a, n in (("src", src, n_ins), ("dst", dst, n_ins), ("w", w, n_ins),
                           ("dsrc", dsrc, n_del), ("ddst", ddst, n_del), ("dw", dw, n_del)):
            if isinstance(a, DeviceArray) and a.n < n:
                raise ValueError("%s holds %d elements, %d needed" % (name, a.n, n))
        st = BatchStats()
        self._check(lib().cytom_update_batch_device(
            self._h, _p(src), _p(dst), _p(w), n_ins, _p(dsrc), _p(ddst), _p(dw), n_del,
            ctypes.byref(st)), "update_device")
        return st

    def update_records(self, records, del_records=None, default_weight=1):
        """One batch from host records in the reference's binary edge layouts:
        8-byte {src, tgt} (file_edge_t / edge_compact_t) or 12-byte
        {src, tgt, float weight} (edge_compact_weighted_t), e.g. slices of an
        np.memmap of the file.  Deletions use the same record size."""
        rec = _records(records)
        drec = _records(del_records) if del_records is not None else rec[:0]
        if len(rec) and len(drec) and drec.dtype.itemsize != rec.dtype.itemsize:
            raise ValueError("insert and delete records differ in size")
        size = rec.dtype.itemsize if len(rec) else drec.dtype.itemsize
        st = BatchStats()
        self._check(lib().cytom_update_batch_records(
            self._h, rec.ctypes.data if len(rec) else None, size, len(rec), default_weight,
            drec.ctypes.data if len(drec) else None, len(drec), ctypes.byref(st)),
            "update_records")
        return st

    def read_results(self, pid):
        """Vertex values in the results_<n> layout (see cytom_read_results):
        uint32 (BFS/CC), float32 (SSSP) or a (delta_sum, rank, delta) float32
        record array (PRD/PPR), V entries."""
        need = self._check(lib().cytom_read_results(self._h, pid, None, 0), "read_results")
        if self.kinds[pid] in (PRD, PPR):
            out = np.empty(self.V, PRD_RECORD)
        elif self.kinds[pid] == SSSP:
            out = np.empty(self.V, np.float32)
        else:
            out = np.empty(self.V, np.uint32)
        assert out.nbytes == need
        self._check(lib().cytom_read_results(self._h, pid, out.ctypes.data, out.nbytes),
                    "read_results")
        return out

    def execute(self, pid):
        st = RunStats()
        self._check(lib().cytom_execute(self._h, pid, ctypes.byref(st)), "execute")
        return st

    def ppr_adjoint(self, pid, dL_dr, tol=0.0, max_iterations=400):
        """Reverse mode of a registered PPR program on the current graph:
        returns (dL/dt float32[V], src, dst, dL/dw float32[E] per live edge in
        export_edges order, sweeps).  dL_dr: float32[V] (host or DeviceArray).
        The defaults run 400 sweeps (0.85^400: FP32 convergence)."""
        g = dL_dr if isinstance(dL_dr, DeviceArray) else \
            DeviceArray.from_host(np.ascontiguousarray(dL_dr, np.float32))
        E = self.live_edge_count()
        dt = DeviceArray(self.V, np.float32)
        s, d, dw = (DeviceArray(max(E, 1), np.uint32), DeviceArray(max(E, 1), np.uint32),
                    DeviceArray(max(E, 1), np.float32))
        it = ctypes.c_uint32()
        n = self._check(lib().cytom_ppr_program_adjoint(
            self._h, pid, _p(g), tol, max_iterations, _p(dt), _p(s), _p(d), _p(dw), max(E, 1),
            ctypes.byref(it)), "ppr_adjoint")
        return dt.to_host(), s.to_host()[:n], d.to_host()[:n], dw.to_host()[:n], it.value

    def execute_many(self, pids):
        """Runs every program in pids on the current graph, concurrently (one
        stream each); returns [RunStats per program]."""
        ids = (ctypes.c_int * max(1, len(pids)))(*pids)
        rs = (RunStats * max(1, len(pids)))()
        self._check(lib().cytom_execute_many(self._h, ctypes.cast(ids, ctypes.c_void_p), len(pids),
                                             ctypes.cast(rs, ctypes.c_void_p)), "execute_many")
        return [rs[k] for k in range(len(pids))]

    def step(self, pids, src, dst, w, n_ins, dsrc=None, ddst=None, dw=None, n_del=0):
        """One update batch from device arrays (as update_device) followed by
        a run of every program in pids, with one host synchronisation when
        the runs converge within their first iteration group.  Returns
        (BatchStats, [RunStats per program])."""
        for name, a, n in (("src", src, n_ins), ("dst", dst, n_ins), ("w", w, n_ins),
                           ("dsrc", dsrc, n_del), ("ddst", ddst, n_del), ("dw", dw, n_del)):
            if isinstance(a, DeviceArray) and a.n < n:
                raise ValueError("%s holds %d elements, %d needed" % (name, a.n, n))
        ids = (ctypes.c_int * max(1, len(pids)))(*pids)
        rs = (RunStats * max(1, len(pids)))()
        st = BatchStats()
        self._check(lib().cytom_step(
            self._h, _p(src), _p(dst), _p(w), n_ins, _p(dsrc), _p(ddst), _p(dw), n_del,
            ctypes.cast(ids, ctypes.c_void_p), len(pids), ctypes.byref(st),
            ctypes.cast(rs, ctypes.c_void_p)), "step")
        return st, [rs[k] for k in range(len(pids))]

    def batched_execute(self, queries):
        """queries: list of (kind, sources); kind BFS/SSSP/PPR or its name.
        Returns the aggregate RunStats; read results with read_query(i)."""
        keep = []
        arr = (Query * max(1, len(queries)))()
        self._qkinds = []
        for i, (kind, sources) in enumerate(queries):
            if isinstance(kind, str):
                kind = PROGRAMS[kind]
            s = np.ascontiguousarray(np.atleast_1d(sources), dtype=np.uint32)
            keep.append(s)
            arr[i] = Query(kind, len(s), s.ctypes.data)
            self._qkinds.append(kind)
        st = RunStats()
        self._check(lib().cytom_batched_execute(self._h, arr, len(queries), ctypes.byref(st)),
                    "batched_execute")
        return st

    def read_query(self, i):
        dt = np.float32 if self._qkinds[i] == PPR else np.uint32
        out = np.empty(self.V, dt)
        self._check(lib().cytom_read_query(self._h, i, _p(out), out.nbytes), "read_query")
        return out

    def read(self, pid):
        dt = np.float32 if self.kinds[pid] in (PRD, PPR) else np.uint32
        out = np.empty(self.V, dt)
        self._check(lib().cytom_read_values(self._h, pid, _p(out), out.nbytes), "read")
        return out

    def degrees(self):
        o = np.empty(self.V, np.uint32)
        i = np.empty(self.V, np.uint32)
        self._check(lib().cytom_read_degrees(self._h, _p(o), _p(i)), "degrees")
        return o, i

    def set_tuning(self, chunk_slots=0, job_chunks=0, blocks_per_cu=0, source_bands=0,
                   resort_pct=0, resort_max_slots=0, prd_frac_bits=0):
        """Zero keeps a field; resort_pct=RESORT_OFF disables sort-compaction."""
        t = Tuning(chunk_slots, job_chunks, blocks_per_cu, source_bands, resort_pct,
                   resort_max_slots, prd_frac_bits)
        self._check(lib().cytom_set_tuning(self._h, ctypes.byref(t)), "tuning")

    def get_tuning(self):
        t = Tuning()
        self._check(lib().cytom_get_tuning(self._h, ctypes.byref(t)), "tuning")
        return {"chunk_slots": t.chunk_slots, "job_chunks": t.job_chunks,
                "blocks_per_cu": t.blocks_per_cu, "source_bands": t.source_bands,
                "resort_pct": t.resort_pct, "resort_max_slots": t.resort_max_slots,
                "prd_frac_bits": t.prd_frac_bits}

    def set_profiling(self, enable=True, every=1):
        """HIP-event timing of gather launches: every launch (every=1) or
        every every-th iteration of each run (the phase rotating per run)."""
        self._check(lib().cytom_set_profiling(self._h, int(every) if enable else 0), "profiling")

    def set_step_graphs(self, enable=True):
        """Steps record their enqueue phase into HIP graphs and replay them
        while the launch arguments repeat (see cytom_set_step_graphs)."""
        self._check(lib().cytom_set_step_graphs(self._h, 1 if enable else 0), "step_graphs")

    def set_concurrent_runs(self, enable=True):
        """Steps run their programs concurrently, one stream each (default)."""
        self._check(lib().cytom_set_concurrent_runs(self._h, 1 if enable else 0),
                    "concurrent_runs")

    def step_graph_stats(self):
        """(graphs recorded, graphs replayed, graphs kept) by this context's steps."""
        c = ctypes.c_uint64()
        r = ctypes.c_uint64()
        k = ctypes.c_uint32()
        self._check(lib().cytom_step_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r),
                                                 ctypes.byref(k)), "step_graph_stats")
        return c.value, r.value, k.value

    def kernel_profile(self, kind):
        """(gather ms, launches, algorithmic bytes) accumulated for a program kind."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        b = ctypes.c_double()
        self._check(lib().cytom_kernel_profile(self._h, kind, ctypes.byref(ms), ctypes.byref(n),
                                               ctypes.byref(b)), "kernel_profile")
        return ms.value, n.value, b.value

    def live_edge_count(self):
        return self._check(lib().cytom_export_edges(self._h, None, None, None, 0), "export")

    def export_edges(self, weights=True):
        """Live edges (src, dst, w); w is None with weights=False."""
        n = self._check(lib().cytom_export_edges(self._h, None, None, None, 0), "export")
        s = np.empty(n, np.uint32)
        d = np.empty(n, np.uint32)
        w = np.empty(n, np.uint32) if weights else None
        m = self._check(lib().cytom_export_edges(self._h, _p(s), _p(d), _p(w), n), "export")
        assert m == n
        return s, d, w
