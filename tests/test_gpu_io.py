"""Record ingest, results export and reference-format jobs on the device,
checked against the oracle through the C-ABI.

* cytom_update_batch_records must build the same store, and the same program
  values, as cytom_update_batch on the same edges (bit-exact: BFS/CC/SSSP are
  min fixpoints and PageRank-delta sums are exact fixed point, so slot order
  cannot change a value);
* cytom_read_results must hold the values of cytom_read_values in the
  results-file layouts;
* a job run of cytom_amd.execute must leave results_<n> files equal to the
  oracle after every batch, and an edge log that replays into the inserted
  graph.
"""
import collections
import json
import os

import numpy as np
import pytest

import cytom_amd as C
import oracle as O
from cytom_amd import execute, formats, jobs, launcher

pytestmark = pytest.mark.gpu

INF = 0xFFFFFFFF
FLT_MAX = np.finfo(np.float32).max
PRD_L1_TOL = 1e-6


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert C.device_count() > 0, "no HIP device visible: the GPU tests need an MI355X"


def l1n(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def multiset(s, d, w):
    return collections.Counter(zip(s.tolist(), d.tolist(), w.tolist()))


def _records(s, d, w=None):
    if w is None:
        r = np.empty(len(s), formats.FILE_EDGE)
    else:
        r = np.empty(len(s), formats.EDGE_COMPACT_WEIGHTED)
        r["weight"] = w.astype(np.float32)
    r["src"] = s
    r["tgt"] = d
    return r


def _register_all(g):
    return {n: g.register(k, 0) for n, k in
            (("bfs", C.BFS), ("cc", C.CC), ("sssp", C.SSSP), ("prd", C.PRD))}


def test_record_ingest_equals_array_ingest():
    scale, seed, vpt = 14, 7, 1024
    V = 1 << scale
    st = O.Stream(seed)
    ga = C.Graph(V, vertices_per_tile=vpt, weighted=True)  # u32 arrays
    gr = C.Graph(V, vertices_per_tile=vpt, weighted=True)  # 12-byte weighted records
    gu = C.Graph(V, vertices_per_tile=vpt, weighted=True)  # 8-byte records, weight 3
    pa, pr, pu = _register_all(ga), _register_all(gr), _register_all(gu)
    # The first batch spans several 8 MiB staging chunks; the last one only deletes.
    for b, (n_ins, n_del) in enumerate([(1_200_000, 0), (5000, 40000), (0, 3000), (777, 1)]):
        print("record ingest batch", b, n_ins, n_del, flush=True)
        first, didx = st.next(n_ins, n_del)
        s, d, w, ds, dd, dw = O.batch_edges(scale, seed, first, n_ins, didx, True)
        ba = ga.update(s, d, w, ds, dd, dw)
        br = gr.update_records(_records(s, d, w), _records(ds, dd, dw))
        bu = gu.update_records(_records(s, d), _records(ds, dd), default_weight=3)
        assert (ba.deleted, ba.live_edges) == (br.deleted, br.live_edges)
        assert ba.live_edges == bu.live_edges
        sa, da, wa = ga.export_edges()
        assert multiset(sa, da, wa) == multiset(*gr.export_edges())
        su, du, wu = gu.export_edges()
        assert np.all(wu == 3)
        assert multiset(sa, da, np.full(len(sa), 3)) == multiset(su, du, wu)
        for n in ("bfs", "cc", "sssp", "prd"):
            ga.execute(pa[n])
            gr.execute(pr[n])
            gu.execute(pu[n])
            va, vr, vu = ga.read(pa[n]), gr.read(pr[n]), gu.read(pu[n])
            assert np.array_equal(va, vr), (b, n)
            if n == "sssp":
                hops = gu.read(pu["bfs"]).astype(np.uint64)
                exp = np.where(hops == INF, INF, 3 * hops).astype(np.uint32)
                assert np.array_equal(vu, exp), b
            else:
                assert np.array_equal(va, vu), (b, n)
    # invalid records change nothing
    live = ga.live_edge_count()
    bad = _records(np.array([1, 2]), np.array([3, 4]), np.array([1.0, 0.5]))
    with pytest.raises(C.CytomError):
        gr.update_records(bad)
    neg = _records(np.array([-1], np.int64), np.array([2]))
    with pytest.raises(C.CytomError):
        gu.update_records(neg, default_weight=3)
    with pytest.raises(C.CytomError):
        gu.update_records(_records(np.array([1]), np.array([2])), default_weight=0)
    assert gr.live_edge_count() == live and gu.live_edge_count() == live


def test_read_results_layouts():
    scale, seed = 12, 4
    V = 1 << scale
    st = O.Stream(seed)
    g = C.Graph(V, vertices_per_tile=256, weighted=True)
    p = _register_all(g)
    ppr = g.register(C.PPR, 5)
    for _ in range(3):
        first, didx = st.next(20000, 2000)
        s, d, w, ds, dd, dw = O.batch_edges(scale, seed, first, 20000, didx, True)
        g.update(s, d, w, ds, dd, dw)
    for n in p:
        g.execute(p[n])
    g.execute(ppr)
    for n in ("bfs", "cc"):
        r = g.read_results(p[n])
        assert r.dtype == np.uint32 and np.array_equal(r, g.read(p[n]))
    r = g.read_results(p["sssp"])
    v = g.read(p["sssp"])
    assert r.dtype == np.float32 and (v == INF).any()
    assert np.array_equal(r, np.where(v == INF, FLT_MAX, v.astype(np.float32)))
    for pid in (p["prd"], ppr):
        r = g.read_results(pid)
        assert r.dtype == formats.PRD_VERTEX and len(r) == V
        assert np.array_equal(r["rank"], g.read(pid))
        assert np.all(np.isfinite(r["delta_sum"])) and np.all(np.isfinite(r["delta"]))


def _job(**kw):
    v = {
        "max-vertices": 4096, "vertices-per-tile": 65536, "batch-insertion-count": 8192,
        "meta-tile-manager-count": 6, "count-edges": 40000, "edge-inserters-count": 6,
        "use-dynamic-compaction": 0, "count-parallel-compaction": 48, "enable-edge-apis": 1,
        "count-algorithm-executors": 24, "algorithm-tile-batching": 256,
        "algorithm-iterations": 100, "path-perf-events": "/tmp/none",
        "enable-perf-event-collection": 0, "enable-static-algorithm-before-compaction": 0,
        "enable-static-algorithm-after-compaction": 0, "enable-rewrite-ids": 0,
        "traversal": "row-first", "generator": "binary", "binary-input-file": "~",
        "algorithm": "bfs", "deletion-percentage": 0.05, "enable-interactive-algorithm": 1,
        "enable-lazy-algorithm": 0, "count-algorithm-appliers": 24,
        "enable-adaptive-scheduling": 1, "tile-distribution-strategy": "tile-distributor",
        "path-to-edges-output": "/tmp/none", "enable-writing-edges": 0,
        "enable-in-memory-ingestion": 0, "td-count-meta-tiles-per-dimension": 1,
        "algo-start-id": 1, "enable-reading-edges": 0, "path-to-edges-input": "/tmp/none",
        "enable-algorithm-reexecution": 0, "path-to-results": "/tmp/none",
        "enable-write-results": 1, "delta": 0.01, "start-algorithm-count": 0,
        "output-detailed-tile-stats": 0}
    v.update(kw)
    return jobs.job_argv(v)


def test_binary_job_results_and_edge_log_replay(tmp_path, capsys):
    scale, V, E, batch = 12, 4096, 40000, 8192
    s, d = O.rmat_edges(scale, 3, 0, E)
    e = np.empty(E, formats.FILE_EDGE)
    e["src"], e["tgt"] = s, d
    gfile = tmp_path / "g.bin"
    e.tofile(str(gfile))
    logs = tmp_path / "edges"
    algos = ["bfs", "cc", "sssp", "pagerank-delta"]
    argv = []
    for a in algos:
        argv += _job(**{"algorithm": a, "binary-input-file": str(gfile),
                        "path-to-results": str(tmp_path / ("res_" + a)),
                        "enable-writing-edges": 1, "path-to-edges-output": str(logs)})
    print("binary job: 4 programs, %d edges in batches of %d" % (E, batch), flush=True)
    assert execute.main(argv) == 0
    summaries = [l for l in capsys.readouterr().out.splitlines() if l.startswith("{")]
    assert len(summaries) == 4

    # oracle replay of the same batches and deletions
    og = O.Graph(V, weighted=True, max_iterations=100, delta=0.01)
    kinds = {"bfs": O.BFS, "cc": O.CC, "sssp": O.SSSP, "pagerank-delta": O.PRD}
    oid = {a: og.register(kinds[a], 1) for a in algos}
    st = O.Stream(execute.RMAT_SEED)
    n_del = int(batch * 0.05)
    done, b, deleted = 0, 0, 0
    while done < E:
        n = min(batch, E - done)
        first, didx = st.next(n, n_del)
        ins, dels = e[first:first + n], e[didx.astype(np.int64)]
        deleted += og.update(ins["src"].astype(np.uint32), ins["tgt"].astype(np.uint32),
                             np.ones(n, np.uint32), dels["src"].astype(np.uint32),
                             dels["tgt"].astype(np.uint32), None)
        for a in algos:
            og.execute(oid[a])
            ov = og.read(oid[a])
            got = formats.read_results(str(tmp_path / ("res_" + a) / ("results_%d" % b)), a)
            if a == "pagerank-delta":
                assert l1n(got["rank"], ov) <= PRD_L1_TOL, b
            elif a == "sssp":
                assert np.array_equal(got, np.where(ov == INF, FLT_MAX, ov.astype(np.float32)))
            else:
                assert np.array_equal(got, ov), (a, b)
        done += n
        b += 1
    import json
    for line in summaries:
        sm = json.loads(line)
        assert sm["inserted"] == E and sm["batches"] == b and sm["results_written"] == b
        assert sm["deleted"] == deleted

    # the edge log holds every insert in order, weighted (the group ran SSSP)
    log = formats.read_edge_log(str(logs / "edges_0"), weighted=True)
    assert np.array_equal(log["src"], s) and np.array_equal(log["tgt"], d)
    assert np.all(log["weight"] == 1.0)

    # replay: the logs rebuild the inserted graph (deletions are not logged),
    # then the static run computes BFS and SSSP on it
    argv = []
    for a in ("bfs", "sssp"):
        argv += _job(**{"algorithm": a, "binary-input-file": str(gfile),
                        "enable-reading-edges": 1, "path-to-edges-input": str(logs),
                        "enable-interactive-algorithm": 0,
                        "enable-static-algorithm-after-compaction": 1,
                        "path-to-results": str(tmp_path / ("replay_" + a))})
    print("replay job from the edge log", flush=True)
    assert execute.main(argv) == 0
    got = formats.read_results(str(tmp_path / "replay_bfs" / "results_0"), "bfs")
    assert np.array_equal(got, O.check_bfs(V, s, d, 1))
    got = formats.read_results(str(tmp_path / "replay_sssp" / "results_0"), "sssp")
    ov = O.check_sssp(V, s, d, np.ones(E, np.uint32), 1)
    assert np.array_equal(got, np.where(ov == INF, FLT_MAX, ov.astype(np.float32)))


def test_launcher_rmat_jobs_in_process(tmp_path, monkeypatch, capsys):
    monkeypatch.setenv("CYTOM_ROOT", str(tmp_path))
    monkeypatch.setenv("CYTOM_ALGORITHM_STORE", str(tmp_path / "store.json"))
    lines = []
    assert launcher.main(["--algorithm", "bfs,cc", "--graph", "rmat-16", "--max-edges", "65536",
                          "--insertion-batch-size", "16384", "--enable-writing-results",
                          "--detailed-tile-stats", "--concurrent", "--print-only"],
                         out=lines.append) == 0
    (line,) = lines
    words = line.split()
    print("launcher job: rmat-16, 65536 edges, bfs + cc", flush=True)
    assert execute.main(words[3:]) == 0
    out = capsys.readouterr().out
    s, d = O.rmat_edges(16, execute.RMAT_SEED, 0, 65536)
    V = 1 << 16
    res = tmp_path / "results-storage"
    # both jobs write to the same path: job i gets its own sub-directory
    bfs = formats.read_results(str(res / "job_0" / "results_3"), "bfs")
    cc = formats.read_results(str(res / "job_1" / "results_3"), "cc")
    assert not (res / "job_0" / "results_4").exists()
    assert np.array_equal(bfs, O.check_bfs(V, s, d, 1))
    assert np.array_equal(cc, O.check_cc(V, s, d))
    # detailed tile stats: "rank count" lines summing to the live edges
    stats = [l.split() for l in out.splitlines() if len(l.split()) == 2 and
             all(x.isdigit() for x in l.split())]
    assert sum(int(c) for _, c in stats) == 65536


def test_launcher_profile_then_concurrent_tasks(tmp_path, monkeypatch):
    # --profile runs each configuration alone and records its device usage;
    # --concurrent then schedules windows of them as child processes that
    # share the GPU, and learns from their measured usage.
    monkeypatch.setenv("CYTOM_ROOT", str(tmp_path))
    monkeypatch.setenv("CYTOM_ALGORITHM_STORE", str(tmp_path / "store.json"))
    state = tmp_path / "scheduler-state.json"
    monkeypatch.setenv("CYTOM_SCHEDULER_STATE", str(state))
    flags = ["--algorithm", "bfs,cc", "--graph", "rmat-16,rmat-20", "--max-edges", "65536",
             "--insertion-batch-size", "16384"]
    for mode, n_tasks in (("--profile", 4), ("--concurrent", None)):
        print("launcher %s: rmat-16 and rmat-20, 65536 edges, bfs + cc" % mode, flush=True)
        out = []
        assert launcher.main(flags + [mode], out=out.append) == 0, "\n".join(out)
        done = [o for o in out if o.startswith("Task ") and "has finished" in o]
        assert all(o.endswith("exit 0)") for o in done)
        if n_tasks:
            assert len(done) == n_tasks
        summaries = [json.loads(o) for o in out if o.startswith("{")]
        assert len(summaries) == 4 and all(s["executions"] == 4 for s in summaries)
        assert {s["algorithm"] for s in summaries} == {"bfs", "cc"}
    st = json.loads(state.read_text())
    prof = st["profile_data"]
    assert sorted(prof) == ["rmat-16", "rmat-20"]
    for g in prof.values():
        assert sorted(g) == ["bfs", "cc"]
        assert all(0.0 < v["gpu_usage"] <= 100.0 and v["running_time"] > 0 for v in g.values())
    assert len(st["model_X"]) >= 1 and len(st["model_X"]) == len(st["model_y"])
    # one log per task, in a directory per launcher run
    runs = sorted(os.listdir(tmp_path / "logs"))
    assert len(runs) == 2 and len(os.listdir(tmp_path / "logs" / runs[0])) == 4
