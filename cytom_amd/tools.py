"""Offline result and graph-file tools with the reference's command lines.

    python -m cytom_amd.tools sort_results --input-results F --output-results G
                                           --sort-ascending 0|1 --algorithm A
    python -m cytom_amd.tools results_diff --result-1 F --result-2 G
                                           --output-file O --count-comparisons N
    python -m cytom_amd.tools kendall_tau --input O --delta D
    python -m cytom_amd.tools convert_graph --input TEXT --output BIN
    python -m cytom_amd.tools rewrite_ids --input-graph BIN --output-graph BIN2

They restate tools/main_sort_results.cc, tools/main_results_diff.cc,
tools/kendall-tau-computation.py, scripts/convert_graph.py and
tools/main-rewrite-ids.cc.  Pure host code over the files the engine writes
(cytom_amd/formats.py).
"""
import argparse
import sys

import numpy as np

from . import formats

# Algorithm names main_sort_results.cc:58-79 accepts, and the ones it sorts.
SORT_NAMES = {"pagerank": "pagerank", "pagerankdelta": "pagerank", "pagerank-delta": "pagerank",
              "connected-components": None, "connected-components-async": None, "bfs": None,
              "bfs-async": None, "sssp": "sssp", "sssp-async": None}
# PageRank values at or below this are dropped (main_sort_results.cc:107-109).
RANK_CUTOFF = 0.1500001


def log(msg):
    print(msg, flush=True)


def sort_values(values, ids, ascending):
    """vertex_value_type_t records ordered by value; equal values keep
    ascending ids in both directions (a stable sort over the forward or the
    reversed range, as std::sort over begin/end or rbegin/rend orders them)."""
    if ascending:
        order = np.argsort(values, kind="stable")
    else:
        order = np.lexsort((ids, -values.astype(np.float64)))
    out = np.empty(len(order), formats.VERTEX_VALUE)
    out["value"] = values[order]
    out["id"] = ids[order]
    return out


def sorted_results(path, algorithm, ascending):
    """main_sort_results.cc:96-140, 170-184."""
    kind = SORT_NAMES[algorithm.lower()]
    if kind == "pagerank":
        rec = np.fromfile(path, dtype=formats.PRD_VERTEX)
        ids = np.arange(len(rec), dtype=np.uint32)
        keep = rec["rank"].astype(np.float64) > RANK_CUTOFF
        return sort_values(rec["rank"][keep], ids[keep], ascending)
    if kind == "sssp":
        vals = np.fromfile(path, dtype="<f4")
        return sort_values(vals, np.arange(len(vals), dtype=np.uint32), ascending)
    raise NotImplementedError("algorithm not implemented: %s" % algorithm)


def results_diff_pairs(lhs, rhs, count):
    """The (lhs id, rhs id) pairs of the first `count` positions
    (main_results_diff.cc:127-160); positions past the shorter file are not
    written."""
    n = min(int(count), len(lhs), len(rhs))
    return np.stack([lhs["id"][:n], rhs["id"][:n]], axis=1)


def kendall_tau(pairs):
    """(tau + 1) / 2 and the p-value of the two rank columns
    (kendall-tau-computation.py:7-19)."""
    from scipy import stats

    a = np.asarray([p[0] for p in pairs])
    b = np.asarray([p[1] for p in pairs])
    tau, p = stats.kendalltau(a, b)
    return (tau + 1.0) / 2.0, p


def _need(args, names):
    missing = [n for n in names if getattr(args, n) is None]
    if missing:
        log("Wrong number of arguments, %d, expected count: %d"
            % (len(names) - len(missing), len(names)))
        return False
    return True


def main_sort_results(argv):
    ap = argparse.ArgumentParser(prog="sort_results")
    ap.add_argument("--input-results")
    ap.add_argument("--output-results")
    ap.add_argument("--sort-ascending", type=int)
    ap.add_argument("--algorithm")
    a = ap.parse_args(argv)
    if not _need(a, ["input_results", "output_results", "sort_ascending", "algorithm"]):
        return 1
    if a.algorithm.lower() not in SORT_NAMES:
        log("Wrong option given for algorithm: %s" % a.algorithm)
        return 1
    try:
        out = sorted_results(a.input_results, a.algorithm, a.sort_ascending == 1)
    except NotImplementedError as e:
        log(str(e))
        return 1
    log("Value size: %d" % len(out))
    for i in range(min(10, len(out))):
        log("Value at %d: %f, %u" % (i, out["value"][i], out["id"][i]))
    with open(a.output_results, "wb") as f:
        out.tofile(f)
    return 0


def main_results_diff(argv):
    ap = argparse.ArgumentParser(prog="results_diff")
    ap.add_argument("--result-1")
    ap.add_argument("--result-2")
    ap.add_argument("--output-file")
    ap.add_argument("--count-comparisons", type=int)
    a = ap.parse_args(argv)
    if not _need(a, ["result_1", "result_2", "output_file", "count_comparisons"]):
        return 1
    pairs = results_diff_pairs(formats.read_vertex_values(a.result_1),
                               formats.read_vertex_values(a.result_2), a.count_comparisons)
    with open(a.output_file, "w") as f:
        for l_id, r_id in pairs:
            f.write("%d %d\n" % (l_id, r_id))
    return 0


def main_kendall_tau(argv):
    ap = argparse.ArgumentParser(prog="kendall_tau")
    ap.add_argument("--input")
    ap.add_argument("--delta")
    a = ap.parse_args(argv)
    pairs = []
    with open(a.input) as f:
        for line in f:
            s = line.split()
            pairs.append((int(s[0]), int(s[1])))
    t, p = kendall_tau(pairs)
    print("%s %f %f" % (a.delta, t, p))
    return 0


def main_convert_graph(argv):
    ap = argparse.ArgumentParser(prog="convert_graph")
    ap.add_argument("--input")
    ap.add_argument("--output")
    a = ap.parse_args(argv)
    n = formats.convert_text(a.input, a.output)
    print("Edges read: %d" % n)
    return 0


def main_rewrite_ids(argv):
    ap = argparse.ArgumentParser(prog="rewrite_ids")
    ap.add_argument("--input-graph")
    ap.add_argument("--output-graph")
    a = ap.parse_args(argv)
    if not _need(a, ["input_graph", "output_graph"]):
        return 1
    edges = formats.open_binary_graph(a.input_graph)
    new, n_ids = formats.rewrite_ids(edges)
    log("Unique vertices: %d" % n_ids)
    log("Count edges: %d" % len(edges))
    with open(a.output_graph, "wb") as f:
        new.tofile(f)
    return 0


TOOLS = {"sort_results": main_sort_results, "results_diff": main_results_diff,
         "kendall_tau": main_kendall_tau, "convert_graph": main_convert_graph,
         "rewrite_ids": main_rewrite_ids}


def main(argv=None):
    argv = sys.argv[1:] if argv is None else list(argv)
    if not argv or argv[0] not in TOOLS:
        print("usage: python -m cytom_amd.tools {%s} [options]" % ",".join(sorted(TOOLS)))
        return 2
    return TOOLS[argv[0]](argv[1:])


if __name__ == "__main__":
    sys.exit(main())
