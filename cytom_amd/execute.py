"""Runs jobs in the reference's 80-argument format on the device engine.

    python -m cytom_amd.execute --max-vertices ... (80 arguments per job) ...

The counterpart of the per-task executable that scripts/run.py links from the
algorithm store and starts with the jobs' argument lists (run.py:723-791
execute_task; core/jobs.h; main/execute.h).  Per job the stream follows
EdgeInserter::run (edge-inserter/edge-inserter.h:53-194):

* the edge source is the first job's generator: RMAT edges from the
  counter-based device generator (``rmat``, max-vertices = 2^scale) or the
  file_edge_t records of ``binary-input-file`` (core/jobs.h:49-75), up to
  count-edges edges, in batches of batch-insertion-count;
* with enable-reading-edges the ``edges_<id>`` logs in path-to-edges-input are
  inserted first and the source continues after them (execute.h:454-462);
* each batch deletes batch-insertion-count x deletion-percentage live edges
  drawn uniformly (meta-tile-manager.h:200-228);
* after each batch, once start-algorithm-count edges are in and
  enable-interactive-algorithm is set, every job's program runs to
  convergence (re-initialised first with enable-algorithm-reexecution, or
  when enable-edge-apis is 0), and with enable-write-results its values go to
  ``results_<n>`` in path-to-results (results-writer.h:46-81);
* with enable-writing-edges the inserted edges are appended to
  ``edges_0`` in path-to-edges-output (edges-writer.h:114-185).

When several jobs of one process write results to the same path-to-results,
job i writes to ``<path>/job_<i>/`` instead, so their ``results_<n>`` files do
not overwrite each other.

Jobs whose graph settings agree share one tile store in HBM and run as
programs of it; other jobs get their own store, one after the other.  Keys
that size CPU thread pools or pick CPU tile schedules (meta-tile-manager-count,
edge-inserters-count, count-parallel-compaction, count-algorithm-executors,
count-algorithm-appliers, algorithm-tile-batching, traversal,
tile-distribution-strategy, td-count-meta-tiles-per-dimension,
use-dynamic-compaction, enable-lazy-algorithm, enable-adaptive-scheduling,
path-perf-events, enable-perf-event-collection) are validated and have no
device counterpart.
"""
import json
import os
import sys
import time

import numpy as np

from . import (SSSP, CytomError, DeviceArray, Graph, Stream, rmat_gather, rmat_generate,
               synchronize)
from . import formats
from .algorithm_store import KIND_NAMES, AlgorithmStore
from .jobs import JobError, parse_jobs

RMAT_SEED = 2
MAX_VPT = 16384  # largest tile the engine's 16-bit local ids allow
SSSP_FILE_WEIGHT = 1  # weight of unweighted file records in a weighted store

# Jobs that agree on these keys share one tile store.
GROUP_KEYS = ("max-vertices", "vertices-per-tile", "batch-insertion-count", "count-edges",
              "deletion-percentage", "algorithm-iterations", "delta", "start-algorithm-count",
              "enable-interactive-algorithm", "enable-reading-edges", "path-to-edges-input",
              "enable-writing-edges", "path-to-edges-output", "enable-rewrite-ids",
              "enable-in-memory-ingestion")


def group_jobs(jobs):
    groups = {}
    for i, j in enumerate(jobs):
        groups.setdefault(tuple(j[k] for k in GROUP_KEYS), []).append(i)
    return list(groups.values())


class _JobRun:
    def __init__(self, index, job, kind, pid, results_dir):
        self.index, self.job, self.kind, self.pid = index, job, kind, pid
        self.results_dir = results_dir
        self.executions = 0
        self.iterations = 0
        self.te = 0
        self.ms = 0.0
        self.results = 0

    def summary(self, acc):
        return {"job": self.index, "algorithm": self.job["algorithm"],
                "program": KIND_NAMES[self.kind], "batches": acc["batches"],
                "inserted": acc["inserted"], "deleted": acc["deleted"],
                "live_edges": acc["live"], "ingest_ms": round(acc["ingest_ms"], 3),
                "executions": self.executions, "iterations": self.iterations,
                "traversed_edges": self.te, "compute_ms": round(self.ms, 3),
                "gteps": self.te / (self.ms * 1e-3) / 1e9 if self.ms > 0 else 0.0,
                "results_written": self.results}


def _tile_stats(g, out):
    """Edge counts of the non-empty tiles, largest first (execute.h:365-381)."""
    s, d, _ = g.export_edges()
    t = (d.astype(np.int64) // g.vpt) * g.T + s.astype(np.int64) // g.vpt
    counts = np.sort(np.bincount(t))[::-1]
    for i, c in enumerate(counts[counts > 0]):
        out("%d %d" % (i, c))


def _source(job0):
    """Edge source of the process: the first job's generator (jobs.h:168-244)."""
    gen = job0["generator"]
    if gen == "binary":
        edges = formats.open_binary_graph(job0["binary-input-file"])
        if job0["enable-rewrite-ids"]:
            edges, _ = formats.rewrite_ids(edges)
        elif job0["enable-in-memory-ingestion"]:
            edges = np.array(edges)
        return gen, edges
    return gen, None


def results_dirs(jobs):
    """path-to-results per job, split per job where writers would collide."""
    users = {}
    for j in jobs:
        if j["enable-write-results"]:
            users[j["path-to-results"]] = users.get(j["path-to-results"], 0) + 1
    return [os.path.join(j["path-to-results"], "job_%d" % i)
            if users.get(j["path-to-results"], 0) > 1 else j["path-to-results"]
            for i, j in enumerate(jobs)]


def run_group(jobs, idxs, kinds, source, out, seed=RMAT_SEED):
    j0 = jobs[idxs[0]]
    V = j0["max-vertices"]
    gen, edges = source
    weighted = any(kinds[i] == SSSP for i in idxs)
    if gen == "rmat":
        scale = V.bit_length() - 1
        if V != 1 << scale:
            raise JobError("the rmat generator needs a power-of-two max-vertices, not %d" % V)
        total = j0["count-edges"]
    else:
        total = min(len(edges), j0["count-edges"])
    batch = j0["batch-insertion-count"]
    if batch == 0:
        raise JobError("batch-insertion-count must be positive")
    vpt = j0["vertices-per-tile"]
    vpt = vpt if 64 <= vpt <= MAX_VPT and vpt & (vpt - 1) == 0 else 0
    g = Graph(V, vertices_per_tile=vpt, weighted=weighted,
              max_iterations=j0["algorithm-iterations"], delta=j0["delta"],
              edge_capacity=max(1 << 22, total + total // 4))
    runs = []
    rdirs = results_dirs(jobs)
    for i in idxs:
        j = jobs[i]
        if j["algo-start-id"] >= V:
            raise JobError("algo-start-id %d is not below max-vertices %d"
                           % (j["algo-start-id"], V))
        runs.append(_JobRun(i, j, kinds[i], g.register(kinds[i], j["algo-start-id"]),
                            rdirs[i]))
    acc = {"batches": 0, "inserted": 0, "deleted": 0, "live": 0, "ingest_ms": 0.0}

    # edge logs of an earlier run (edges-writer.h:31-105)
    offset = 0
    if j0["enable-reading-edges"]:
        for path in formats.edge_log_files(j0["path-to-edges-input"]):
            log = formats.read_edge_log(path, weighted)
            for s in range(0, len(log), batch):
                st = g.update_records(log[s:s + batch], default_weight=SSSP_FILE_WEIGHT)
                acc["ingest_ms"] += st.t_ingest_ms
                acc["live"] = st.live_edges
            offset += len(log)
        out("read %d edges from the edge logs in %s" % (offset, j0["path-to-edges-input"]))
        if offset > total:
            raise JobError("the edge logs hold %d edges, more than count-edges %d"
                           % (offset, total))
        acc["inserted"] = offset
    writer = (formats.EdgeLogWriter(j0["path-to-edges-output"], 0, weighted)
              if j0["enable-writing-edges"] else None)

    stream = Stream(seed)
    n_del = int(batch * j0["deletion-percentage"])
    remaining = total - offset
    t0 = time.perf_counter()
    while remaining > 0:
        n_ins = min(batch, remaining)
        first, didx = stream.next(n_ins, n_del)
        if gen == "rmat":
            s, d, w = rmat_generate(scale, seed, offset + first, n_ins, weights=weighted)
            if len(didx):
                di = DeviceArray.from_host(didx + np.uint64(offset))
                ds, dd, dw = rmat_gather(scale, seed, di, len(didx), weights=weighted)
            else:
                ds = dd = dw = None
            st = g.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
            if writer:
                writer.append(s.to_host(), d.to_host(), w.to_host() if weighted else None)
        else:
            rec = edges[offset + first: offset + first + n_ins]
            drec = edges[offset + didx.astype(np.int64)] if len(didx) else None
            st = g.update_records(rec, drec, default_weight=SSSP_FILE_WEIGHT)
            if writer:
                writer.append(rec["src"], rec["tgt"],
                              np.full(n_ins, SSSP_FILE_WEIGHT, np.float32) if weighted else None)
        remaining -= n_ins
        acc["batches"] += 1
        acc["inserted"] += n_ins
        acc["deleted"] += st.deleted
        acc["live"] = st.live_edges
        acc["ingest_ms"] += st.t_ingest_ms
        line = "batch %d: +%d -%d live %d" % (acc["batches"], n_ins, st.deleted, st.live_edges)
        if j0["enable-interactive-algorithm"] and acc["inserted"] >= j0["start-algorithm-count"]:
            # The jobs sharing this graph run together, concurrently on the device.
            stats = _execute_all(g, runs, [bool(r.job["enable-algorithm-reexecution"]
                                                or not r.job["enable-edge-apis"]) for r in runs])
            for r, rs in zip(runs, stats):
                line += " | job %d %s: %d it, %d edges, %.3f ms" % (
                    r.index, r.job["algorithm"], rs.iterations, rs.traversed_edges,
                    rs.t_compute_ms)
        out(line)
    if writer:
        writer.close()
    for r in runs:
        if (r.job["enable-static-algorithm-before-compaction"]
                or r.job["enable-static-algorithm-after-compaction"]):
            rs = _execute(g, r, reinit=True)
            out("static job %d %s: %d it, %d edges, %.3f ms" % (
                r.index, r.job["algorithm"], rs.iterations, rs.traversed_edges, rs.t_compute_ms))
    synchronize()
    wall = time.perf_counter() - t0
    if j0["output-detailed-tile-stats"]:
        _tile_stats(g, out)
    summaries = [dict(r.summary(acc), wall_s=round(wall, 3)) for r in runs]
    g.close()
    return summaries


def _execute_all(g, runs, reinit):
    for r, ri in zip(runs, reinit):
        if ri:
            g.init(r.pid)
    stats = g.execute_many([r.pid for r in runs])
    for r, rs in zip(runs, stats):
        _account(g, r, rs)
    return stats


def _execute(g, r, reinit):
    if reinit:
        g.init(r.pid)
    rs = g.execute(r.pid)
    _account(g, r, rs)
    return rs


def _account(g, r, rs):
    r.executions += 1
    r.iterations += rs.iterations
    r.te += rs.traversed_edges
    r.ms += rs.t_compute_ms
    if r.job["enable-write-results"]:
        vals = g.read_results(r.pid)[:r.job["max-vertices"]]
        formats.write_results(r.results_dir, r.results, vals)
        r.results += 1


def run_jobs(jobs, store=None, out=print, seed=RMAT_SEED):
    """Runs parsed jobs (jobs.parse_jobs); returns one summary dict per job."""
    store = store or AlgorithmStore(out=out)
    kinds = []
    for j in jobs:
        k = store.resolve(j["algorithm"])
        if k is None:
            raise JobError("algorithm '%s' is neither in the algorithm store nor built in"
                           % j["algorithm"])
        kinds.append(k)
    source = _source(jobs[0])
    summaries = [None] * len(jobs)
    for store, idxs in enumerate(group_jobs(jobs)):
        for s in run_group(jobs, idxs, kinds, source, out, seed):
            s["store"] = store  # jobs with the same index shared one tile store
            summaries[s["job"]] = s
    return summaries


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    try:
        jobs = parse_jobs(argv)
        summaries = run_jobs(jobs)
    except (JobError, CytomError) as e:
        print("error: %s" % e, file=sys.stderr, flush=True)
        return 1
    for s in summaries:
        print(json.dumps(s), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
