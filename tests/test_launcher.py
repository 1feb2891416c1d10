"""Host side of the launcher, job format, algorithm store, file formats and
result tools (no device needed).

The expected job argument lists below are written out from
scripts/run.py:56-148 and scripts/config.py by hand, value by value, so a
change of key order, default or derived value in the launcher fails here.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from cytom_amd import algorithm_store, formats, jobs, launcher, tools


def _args(*argv):
    return launcher.build_parser().parse_args(list(argv))


@pytest.fixture()
def roots(tmp_path, monkeypatch):
    monkeypatch.setenv("CYTOM_ROOT", str(tmp_path / "root"))
    monkeypatch.setenv("CYTOM_DATA_ROOT", str(tmp_path / "data"))
    monkeypatch.setenv("CYTOM_ALGORITHM_STORE", str(tmp_path / "store.json"))
    monkeypatch.setenv("CYTOM_SCHEDULER_STATE", str(tmp_path / "scheduler-state.json"))
    return tmp_path


# ------------------------------------------------------------------ job format
def test_job_keys_follow_execute_h_positions():
    # execute.h reads these string options by argv position.
    pos = {k: 2 * i + 2 for i, (k, _) in enumerate(jobs.JOB_KEYS)}
    assert pos["path-perf-events"] == 26
    assert pos["traversal"] == 36
    assert pos["generator"] == 38
    assert pos["binary-input-file"] == 40
    assert pos["tile-distribution-strategy"] == 54
    assert pos["path-to-edges-output"] == 56
    assert pos["path-to-edges-input"] == 68
    assert pos["path-to-results"] == 72
    assert len(jobs.JOB_KEYS) * 2 == jobs.NUM_OF_ARGS == 80


def test_rmat22_bfs_job_golden(roots):
    a = _args("--algorithm", "bfs", "--graph", "rmat-22", "--sequential")
    (c,) = launcher.configurations(a, algorithm_store.AlgorithmStore())
    argv = jobs.job_argv(launcher.job_values(a, c))
    r = str(roots / "root")
    expected = [
        "--max-vertices", "4194304", "--vertices-per-tile", "65536",
        "--batch-insertion-count", "1048576", "--meta-tile-manager-count", "6",
        "--count-edges", "67108864", "--edge-inserters-count", "6",
        "--use-dynamic-compaction", "0", "--count-parallel-compaction", "48",
        "--enable-edge-apis", "1", "--count-algorithm-executors", "24",
        "--algorithm-tile-batching", "256", "--algorithm-iterations", "100",
        "--path-perf-events", r + "/perf_events", "--enable-perf-event-collection", "0",
        "--enable-static-algorithm-before-compaction", "0",
        "--enable-static-algorithm-after-compaction", "0", "--enable-rewrite-ids", "0",
        "--traversal", "row-first", "--generator", "rmat", "--binary-input-file", "~",
        "--algorithm", "bfs", "--deletion-percentage", "0.00",
        "--enable-interactive-algorithm", "1", "--enable-lazy-algorithm", "0",
        "--count-algorithm-appliers", "24", "--enable-adaptive-scheduling", "1",
        "--tile-distribution-strategy", "tile-distributor",
        "--path-to-edges-output", r + "/edges-output", "--enable-writing-edges", "0",
        "--enable-in-memory-ingestion", "0", "--td-count-meta-tiles-per-dimension", "8",
        "--algo-start-id", "1", "--enable-reading-edges", "0",
        "--path-to-edges-input", r + "/edges-input", "--enable-algorithm-reexecution", "0",
        "--path-to-results", r + "/results-storage", "--enable-write-results", "0",
        "--delta", "0.01", "--start-algorithm-count", "0", "--output-detailed-tile-stats", "0",
    ]
    assert argv == expected
    job = jobs.parse_job(argv)
    assert job["max-vertices"] == 1 << 22 and job["delta"] == 0.01
    assert job["generator"] == "rmat" and job["algorithm"] == "bfs"


def test_binary_graph_job_and_flags(roots):
    a = _args("--algorithm", "sssp", "--graph", "livejournal", "--max-edges", "1000",
              "--algorithm-threshold", "0.5", "--enable-writing-results", "--enable-write-edges",
              "--read-edges", "1", "--enable-algorithm-reexecution", "--max-iterations", "7",
              "--algorithm-start-vertex", "3", "--detailed-tile-stats", "--collect-perf-events",
              "--disable-selective-scheduling", "--sequential")
    (c,) = launcher.configurations(a, algorithm_store.AlgorithmStore())
    v = launcher.job_values(a, c)
    assert v["max-vertices"] == 4847571 and v["count-edges"] == 1000
    assert v["generator"] == "binary"
    assert v["binary-input-file"] == str(roots / "data" / "input" / "livejournal" /
                                         "livejournal.bin")
    # 4847571 / 65536 / 4 = 18.49 -> previous_power_of_2 = 2^ceil(log2(18.49) - 1) = 16
    assert v["td-count-meta-tiles-per-dimension"] == 16
    assert v["start-algorithm-count"] == 500
    assert v["enable-write-results"] == "1" and v["enable-writing-edges"] == "1"
    assert v["enable-reading-edges"] == "1" and v["enable-algorithm-reexecution"] == "1"
    assert v["algorithm-iterations"] == 7 and v["algo-start-id"] == 3
    assert v["output-detailed-tile-stats"] == "1"
    assert v["enable-perf-event-collection"] == "1"
    assert v["enable-adaptive-scheduling"] == "0"
    a2 = _args("--algorithm", "bfs", "--graph", "rmat-16", "--algorithm-threshold", "-1",
               "--disable-algorithm", "--sequential")
    (c2,) = launcher.configurations(a2, algorithm_store.AlgorithmStore())
    v2 = launcher.job_values(a2, c2)
    # 2^16 / 2^16 / 4 = 0.25 -> 2^-3, raised to 1
    assert v2["td-count-meta-tiles-per-dimension"] == 1
    assert v2["start-algorithm-count"] == 0
    assert (v2["meta-tile-manager-count"], v2["edge-inserters-count"]) == (24, 10)
    assert v2["enable-interactive-algorithm"] == "0"


def test_configuration_cross_product(roots):
    a = _args("--algorithm", "bfs,cc", "--graph", "rmat-16,rmat-20", "--enable-deletions",
              "--enable-delta", "--runs", "2", "--insertion-batch-size", "1024..4096",
              "--tile-batch-size", "8..16", "--enable-traversals", "--enable-edge-apis",
              "--override-selective-scheduling", "--sequential")
    confs = launcher.configurations(a, algorithm_store.AlgorithmStore())
    # ConfigLists.countConfigs (run.py:378-385)
    assert len(confs) == 2 * 2 * 2 * 3 * 2 * 3 * 2 * 3 * 2 * 10
    # nested-loop order: run outermost, delta innermost
    assert [c["delta"] for c in confs[:10]] == launcher.DELTAS
    assert confs[0]["run"] == 0 and confs[-1]["run"] == 1
    # fields after insertion_batch: tile 2 x traversal 3 x edge-apis 2 x deletions 3 x
    # selective 2 x delta 10 = 720 configurations per batch size
    assert {c["insertion_batch"] for c in confs[:720]} == {1024}
    assert confs[720]["insertion_batch"] == 2048
    assert sorted({c["insertion_batch"] for c in confs}) == [1024, 2048, 4096]
    assert sorted({c["tile_batch"] for c in confs}) == [8, 16]
    assert {c["deletions"] for c in confs} == {"0.01", "0.05", "0.10"}
    a3 = _args("--insertion-batch-size", "5,7", "--algorithm", "bfs", "--graph", "rmat-16",
               "--tile-distribution-strategy", "x", "--sequential")
    c3 = launcher.configurations(a3, algorithm_store.AlgorithmStore())
    assert [c["insertion_batch"] for c in c3][::4] == [5, 7]
    assert [c["tile_distribution_strategy"] for c in c3][:4] == \
        launcher.TILE_DISTRIBUTION_STRATEGIES


def test_print_only_and_concurrent_grouping(roots):
    buf = []
    rc = launcher.main(["--algorithm", "bfs,cc", "--graph", "rmat-16,rmat-20", "--concurrent",
                        "--print-only"], out=buf.append)
    assert rc == 0 and len(buf) == 2  # one process per graph
    for line in buf:
        words = line.split()
        assert words[1:3] == ["-m", "cytom_amd.execute"]
        js = jobs.parse_jobs(words[3:])
        assert [j["algorithm"] for j in js] == ["bfs", "cc"]
    buf = []
    launcher.main(["--algorithm", "bfs,cc", "--graph", "rmat-16", "--sequential", "--print-only"],
                  out=buf.append)
    assert len(buf) == 2 and all(len(b.split()) == 3 + 80 for b in buf)


def test_parse_job_errors():
    good = jobs.job_argv({k: ("rmat" if k == "generator" else "row-first" if k == "traversal"
                              else "static" if k == "tile-distribution-strategy" else "1")
                          for k, _ in jobs.JOB_KEYS})
    jobs.parse_job(good)
    with pytest.raises(jobs.JobError, match="multiple of"):
        jobs.split_jobs(good[:-2])
    swapped = list(good)
    swapped[0], swapped[2] = swapped[2], swapped[0]
    with pytest.raises(jobs.JobError, match="belongs"):
        jobs.parse_job(swapped)
    bad = list(good)
    bad[bad.index("--traversal") + 1] = "zigzag"
    with pytest.raises(jobs.JobError, match="Wrong option given for traversal"):
        jobs.parse_job(bad)
    bad = list(good)
    bad[bad.index("--max-vertices") + 1] = "-5"
    with pytest.raises(jobs.JobError, match="invalid"):
        jobs.parse_job(bad)
    upper = list(good)
    upper[upper.index("--generator") + 1] = "RMAT"
    assert jobs.parse_job(upper)["generator"] == "rmat"
    assert len(jobs.split_jobs(good + good)) == 2


# ------------------------------------------------------------------ algorithm store
def test_algorithm_store_actions(roots):
    src = roots / "mybfs.cc"
    src.write_text("int main_bfs(int argc, char** argv) { return 0; }\n")
    other = roots / "nothing.cc"
    other.write_text("int main(void) { return 0; }\n")
    out = []

    def run(*argv):
        out.clear()
        assert launcher.main(["--algorithm-store"] + list(argv), out=out.append) == 0
        return list(out)

    assert "Successfully add mybfs" in run("--add", "--name", "mybfs", "--path", str(src))[0]
    assert "Successfully overwrite mybfs" in run("--add", "--name", "mybfs", "--path",
                                                 str(src))[0]
    assert "defines no vertex program" in run("--add", "--name", "x", "--path", str(other))[0]
    assert "does not equal" in run("--add", "--name", "a", "b", "--path", str(src))[0]
    run("--add", "--name", "mypr", "--path", "pagerank-delta")
    assert run("--list") == ["mybfs\nmypr"]
    st = algorithm_store.AlgorithmStore()
    assert st.resolve("mybfs") == 0 and st.resolve("mypr") == 3 and st.resolve("sssp") == 2
    assert st.resolve("unknown") is None
    assert "not a valid algorithm name" in run("--delete", "--name", "zzz")[0]
    run("--delete", "--name", "mybfs")
    assert run("--list") == ["mypr"]
    assert "Successfully delete all" in run("--clear")[0]
    assert run("--list") == [""]
    assert "Please specify action" in run()[0]
    # with no --algorithm the launcher takes the store's list (config.ALGORITHMS)
    run("--add", "--name", "b1", "--path", "bfs")
    a = _args("--graph", "rmat-16", "--sequential")
    assert [c["algorithm"] for c in
            launcher.configurations(a, algorithm_store.AlgorithmStore())] == ["b1"]


def test_launcher_reports_failed_task(roots):
    # A job naming an unknown algorithm fails in the child process before any
    # device work; the launcher counts it as failed.
    out = []
    rc = launcher.main(["--algorithm", "no-such-program", "--graph", "rmat-16", "--sequential"],
                       out=out.append)
    assert rc == 1
    assert any("failed: 1" in o for o in out)


# ------------------------------------------------------------------ file formats
def test_convert_text_matches_struct_packing(tmp_path):
    txt = tmp_path / "g.txt"
    txt.write_text("# comment\n1 2\n3\t4\n# x\n-7 2147483647\n")
    out = tmp_path / "g.bin"
    assert formats.convert_text(str(txt), str(out)) == 3
    expect = b"".join(struct.pack("<l", v) for v in (1, 2, 3, 4, -7, 2147483647))
    assert out.read_bytes() == expect
    e = formats.open_binary_graph(str(out))
    assert e["src"].tolist() == [1, 3, -7] and e["tgt"].tolist() == [2, 4, 2147483647]
    txt.write_text("1 4294967296\n")
    with pytest.raises(ValueError):
        formats.convert_text(str(txt), str(out))


def test_rewrite_ids_first_appearance():
    rng = np.random.default_rng(3)
    e = np.empty(500, formats.FILE_EDGE)
    e["src"] = rng.integers(-50, 1000, 500)
    e["tgt"] = rng.integers(-50, 1000, 500)
    new, n = formats.rewrite_ids(e)
    # brute force: the map of main-rewrite-ids.cc:70-105
    m = {}
    exp = []
    for s, t in zip(e["src"].tolist(), e["tgt"].tolist()):
        for x in (s, t):
            if x not in m:
                m[x] = len(m)
        exp.append((m[s], m[t]))
    assert n == len(m)
    assert list(zip(new["src"].tolist(), new["tgt"].tolist())) == exp


def test_rewrite_ids_tool(tmp_path):
    e = np.array([(10, 5), (5, 99), (99, 10)], formats.FILE_EDGE)
    p = tmp_path / "in.bin"
    e.tofile(str(p))
    q = tmp_path / "out.bin"
    assert tools.main(["rewrite_ids", "--input-graph", str(p), "--output-graph", str(q)]) == 0
    r = formats.open_binary_graph(str(q))
    assert r.tolist() == [(0, 1), (1, 2), (2, 0)]


def test_edge_log_round_trip(tmp_path):
    d = tmp_path / "logs"
    with formats.EdgeLogWriter(str(d), 0, weighted=True) as w:
        w.append(np.array([1, 2]), np.array([3, 4]), np.array([5, 6]))
        w.append(np.array([7]), np.array([8]), np.array([9]))
    with formats.EdgeLogWriter(str(d), 2, weighted=False) as w:
        w.append(np.array([1]), np.array([2]))
    (d / "edges_x").write_bytes(b"")
    files = formats.edge_log_files(str(d))
    assert [os.path.basename(f) for f in files] == ["edges_0", "edges_2"]
    a = formats.read_edge_log(files[0], weighted=True)
    assert a.tolist() == [(1, 3, 5.0), (2, 4, 6.0), (7, 8, 9.0)]
    assert os.path.getsize(files[0]) == 3 * 12
    assert formats.read_edge_log(files[1]).tolist() == [(1, 2)]
    assert formats.edge_log_files(str(tmp_path / "none")) == []


def test_results_files(tmp_path):
    v = np.arange(10, dtype=np.uint32)
    p = formats.write_results(str(tmp_path), 3, v)
    assert os.path.basename(p) == "results_3"
    assert np.array_equal(formats.read_results(p, "bfs"), v)
    r = np.zeros(4, formats.PRD_VERTEX)
    r["rank"] = [0.1, 0.2, 0.3, 0.4]
    p = formats.write_results(str(tmp_path), 4, r)
    assert os.path.getsize(p) == 48
    assert np.array_equal(formats.read_results(p, "pagerank-delta")["rank"], r["rank"])


# ------------------------------------------------------------------ tools
def _prd_file(path, ranks):
    r = np.zeros(len(ranks), formats.PRD_VERTEX)
    r["rank"] = ranks
    r.tofile(str(path))


def test_sort_results_pagerank_and_sssp(tmp_path, capsys):
    ranks = np.array([0.15, 0.9, 0.2, 0.1500001, 0.9, 0.5, 0.16], np.float32)
    src = tmp_path / "results_0"
    _prd_file(src, ranks)
    dst = tmp_path / "sorted"
    assert tools.main(["sort_results", "--input-results", str(src), "--output-results", str(dst),
                       "--sort-ascending", "0", "--algorithm", "pagerank-delta"]) == 0
    got = formats.read_vertex_values(str(dst))
    # brute force: keep rank > 0.1500001 (as a double), descending, ties by id
    keep = [(float(x), i) for i, x in enumerate(ranks) if float(x) > 0.1500001]
    exp = sorted(keep, key=lambda t: (-t[0], t[1]))
    assert [(float(v), int(i)) for v, i in got.tolist()] == exp
    assert "Value size: %d" % len(exp) in capsys.readouterr().out
    sssp = np.array([3.0, 1.0, np.finfo(np.float32).max, 1.0, 0.0], np.float32)
    sssp.tofile(str(src))
    assert tools.main(["sort_results", "--input-results", str(src), "--output-results", str(dst),
                       "--sort-ascending", "1", "--algorithm", "SSSP"]) == 0
    got = formats.read_vertex_values(str(dst))
    assert got["id"].tolist() == [4, 1, 3, 0, 2]
    assert tools.main(["sort_results", "--input-results", str(src), "--output-results", str(dst),
                       "--sort-ascending", "1", "--algorithm", "bfs"]) == 1
    assert tools.main(["sort_results", "--input-results", str(src), "--output-results", str(dst),
                       "--sort-ascending", "1", "--algorithm", "nope"]) == 1
    assert tools.main(["sort_results", "--input-results", str(src)]) == 1


def test_results_diff_and_kendall_tau(tmp_path, capsys):
    a = np.zeros(6, formats.VERTEX_VALUE)
    b = np.zeros(5, formats.VERTEX_VALUE)
    a["id"] = [5, 4, 3, 2, 1, 0]
    b["id"] = [5, 3, 4, 1, 2]
    pa, pb = tmp_path / "a", tmp_path / "b"
    a.tofile(str(pa))
    b.tofile(str(pb))
    out = tmp_path / "cmp"
    assert tools.main(["results_diff", "--result-1", str(pa), "--result-2", str(pb),
                       "--output-file", str(out), "--count-comparisons", "4"]) == 0
    assert out.read_text() == "5 5\n4 3\n3 4\n2 1\n"
    assert tools.main(["kendall_tau", "--input", str(out), "--delta", "0.01"]) == 0
    delta, t, p = capsys.readouterr().out.split()
    # brute force tau over the 4 pairs (no ties): concordant - discordant
    xs, ys = [5, 4, 3, 2], [5, 3, 4, 1]
    c = sum(np.sign(xs[i] - xs[j]) * np.sign(ys[i] - ys[j])
            for i in range(4) for j in range(i + 1, 4))
    assert delta == "0.01"
    assert abs(float(t) - (c / 6 + 1) / 2) < 1e-6
    assert 0.0 <= float(p) <= 1.0


def test_tools_cli_entry_point():
    r = subprocess.run([sys.executable, "-m", "cytom_amd.tools"], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 2 and "sort_results" in r.stdout
