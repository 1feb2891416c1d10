"""On-disk formats of the reference, read and written with numpy.

Layouts (all little-endian, no headers):

* binary graph inputs: file_edge_t {int32 src, int32 tgt}
  (util/datatypes.h:41-44), written by scripts/convert_graph.py:8-18 and
  mmap-read by the binary generator (core/jobs.h:52-75);
* edge logs ``edges_<id>``: edge_compact_t {u32 src, u32 tgt} or, for
  weighted graphs, edge_compact_weighted_t {u32 src, u32 tgt, float weight}
  (util/datatypes.h:55-64; tile-manager/edges-writer.h:114-185 writes,
  :31-105 reads them back);
* results ``results_<n>``: the raw VertexType[V] array after the n-th
  algorithm execution (core/results-writer.h:46-81): uint32 for BFS/CC, float
  for SSSP, {float delta_sum, rank, delta} for PageRank-delta;
* sorted results: vertex_value_type_t {float value, u32 id}
  (util/datatypes.h:324-331), written by tools/main_sort_results.cc.
"""
import os
import re

import numpy as np

FILE_EDGE = np.dtype([("src", "<i4"), ("tgt", "<i4")])
EDGE_COMPACT = np.dtype([("src", "<u4"), ("tgt", "<u4")])
EDGE_COMPACT_WEIGHTED = np.dtype([("src", "<u4"), ("tgt", "<u4"), ("weight", "<f4")])
PRD_VERTEX = np.dtype([("delta_sum", "<f4"), ("rank", "<f4"), ("delta", "<f4")])
VERTEX_VALUE = np.dtype([("value", "<f4"), ("id", "<u4")])

_EDGES_RE = re.compile(r"^edges_(\d+)$")


# ---------------------------------------------------------------- graph inputs
def open_binary_graph(path):
    """Read-only memory map of a file_edge_t file (trailing partial record
    ignored, as file_size / sizeof(file_edge_t) in jobs.h:56)."""
    n = os.path.getsize(path) // FILE_EDGE.itemsize
    if n == 0:
        return np.zeros(0, FILE_EDGE)
    return np.memmap(path, dtype=FILE_EDGE, mode="r", shape=(n,))


def convert_text(input_path, output_path):
    """Text edge list -> file_edge_t (scripts/convert_graph.py:8-18): lines
    starting with '#' are skipped, every other line holds exactly two integer
    ids.  Returns the number of edges written."""
    chunk = []
    count = 0
    with open(output_path, "wb") as out:
        with open(input_path) as f:
            for line in f:
                if line.startswith("#"):
                    continue
                src, tgt = line.split()
                chunk.append((int(src), int(tgt)))
                if len(chunk) >= 1 << 20:
                    _pack_i32(chunk).tofile(out)
                    count += len(chunk)
                    chunk = []
        if chunk:
            _pack_i32(chunk).tofile(out)
            count += len(chunk)
    return count


def _pack_i32(pairs):
    a = np.array(pairs, np.int64)
    if a.size and (a.min() < -(1 << 31) or a.max() >= (1 << 31)):
        raise ValueError("vertex id outside the int32 range of file_edge_t")
    return a.astype("<i4")


def rewrite_ids(edges):
    """Dense renumbering in order of first appearance, the source before the
    target of each edge (tools/main-rewrite-ids.cc:70-105).  edges: file_edge_t
    array.  Returns (new file_edge_t array, number of distinct ids)."""
    seq = np.empty(2 * len(edges), np.int64)
    seq[0::2] = edges["src"]
    seq[1::2] = edges["tgt"]
    uniq, first, inv = np.unique(seq, return_index=True, return_inverse=True)
    # rank of each distinct id by its first position in the sequence
    order = np.argsort(first, kind="stable")
    new_of_uniq = np.empty(len(uniq), np.int64)
    new_of_uniq[order] = np.arange(len(uniq))
    mapped = new_of_uniq[inv]
    out = np.empty(len(edges), FILE_EDGE)
    out["src"] = mapped[0::2]
    out["tgt"] = mapped[1::2]
    return out, len(uniq)


# ---------------------------------------------------------------- edge logs
def edge_log_dtype(weighted):
    return EDGE_COMPACT_WEIGHTED if weighted else EDGE_COMPACT


class EdgeLogWriter:
    """Appends inserted edges to ``<dir>/edges_<id>`` (edges-writer.h:114-185;
    the file is opened with O_SYNC there, here it is fsync'ed on close)."""

    def __init__(self, directory, log_id=0, weighted=False):
        os.makedirs(directory, exist_ok=True)
        self.path = os.path.join(directory, "edges_%d" % log_id)
        self.dtype = edge_log_dtype(weighted)
        self._f = open(self.path, "wb")
        self.count = 0

    def append(self, src, dst, w=None):
        rec = np.empty(len(src), self.dtype)
        rec["src"] = src
        rec["tgt"] = dst
        if "weight" in self.dtype.names:
            if w is None:
                raise ValueError("weighted edge log needs weights")
            rec["weight"] = np.asarray(w, np.float32)
        rec.tofile(self._f)
        self.count += len(rec)

    def close(self):
        if self._f:
            self._f.flush()
            os.fsync(self._f.fileno())
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def edge_log_files(directory):
    """``edges_<id>`` files of a directory in id order."""
    if not os.path.isdir(directory):
        return []
    found = []
    for name in os.listdir(directory):
        m = _EDGES_RE.match(name)
        if m:
            found.append((int(m.group(1)), os.path.join(directory, name)))
    return [p for _, p in sorted(found)]


def read_edge_log(path, weighted=False):
    """Memory map of one edge log (count = file size / record size,
    edges-writer.h:86-105)."""
    dt = edge_log_dtype(weighted)
    n = os.path.getsize(path) // dt.itemsize
    if n == 0:
        return np.zeros(0, dt)
    return np.memmap(path, dtype=dt, mode="r", shape=(n,))


# ---------------------------------------------------------------- results
def results_dtype(algorithm):
    """VertexType of a results file by algorithm name."""
    a = algorithm.lower()
    if a in ("pagerank-delta", "pagerankdelta", "pagerank", "ppr"):
        return PRD_VERTEX
    if a in ("sssp", "sssp-async"):
        return np.dtype("<f4")
    return np.dtype("<u4")


def write_results(directory, count, values):
    """``<dir>/results_<count>`` (results-writer.h:54-73)."""
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "results_%d" % count)
    with open(path, "wb") as f:
        np.ascontiguousarray(values).tofile(f)
        f.flush()
        os.fsync(f.fileno())
    return path


def read_results(path, algorithm):
    dt = results_dtype(algorithm)
    return np.fromfile(path, dtype=dt, count=os.path.getsize(path) // dt.itemsize)


def read_vertex_values(path):
    """A sorted-results file of vertex_value_type_t records."""
    return np.fromfile(path, dtype=VERTEX_VALUE,
                       count=os.path.getsize(path) // VERTEX_VALUE.itemsize)
