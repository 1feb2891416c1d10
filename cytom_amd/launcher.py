"""run.py-compatible front end.

    python -m cytom_amd.launcher --algorithm bfs,pagerank-delta --graph rmat-22 --sequential
    python -m cytom_amd.launcher --algorithm-store --add --name mybfs --path examples/bfs.cc

Takes the flag surface of scripts/run.py:941-997, expands it into the same
cross product of configurations (run.py:198-376 GraphConfigGenerator /
ConfigLists), turns each into the 40-key job of run.py:107-148
(SingleGraphConfig.generateArguments) and runs the jobs through
``python -m cytom_amd.execute``:

* ``--sequential`` / ``--profile``: one process per configuration, one after
  the other, each recording its device usage as that algorithm's profile on
  that graph (run.py:905-914 profile_all; cytom_amd/scheduler.py);
* ``--concurrent``: the device scheduler of cytom_amd/scheduler.py
  (run.py:415-681): per graph, windows of configurations run as the jobs of
  one process, sharing its tile store, and several such processes share the
  GPU while their predicted device usage and memory fit;
* ``--print-only``: prints the job command lines instead of running them;
* ``--algorithm-store`` with ``--add/--delete/--clear/--list``: the registry
  of cytom_amd/algorithm_store.py (run.py:683-759).

The graph tables and defaults restate scripts/config.py (sizes of the public
graphs, RMAT scales, default batch sizes and paths); files are looked up below
CYTOM_DATA_ROOT (default /data/graph/), outputs go below CYTOM_ROOT (default
the working directory).
"""
import argparse
import math
import os
import shlex
import sys
import time
from collections import OrderedDict

from .algorithm_store import AlgorithmStore
from .jobs import job_argv
from .scheduler import Algo, Scheduler, task_command

# ---------------------------------------------------------------- scripts/config.py
GRAPH_SETTINGS_DELIM = {
    "twitter-small": {"vertices": 1334392, "edges": 46180342},
    "twitter-full": {"vertices": 41652230, "edges": 1468365182},
    "friendster": {"vertices": 65608366, "edges": 1806067135},
    "livejournal": {"vertices": 4847571, "edges": 68993773},
    "orkut": {"vertices": 3072441, "edges": 117185083},
    "uk2007": {"vertices": 105896555, "edges": 3738733599},
}
GRAPH_SETTINGS_RMAT = {
    "rmat-16": {"vertices": 2 ** 16, "edges": 2 ** 20},
    "rmat-20": {"vertices": 2 ** 20, "edges": 2 ** 24},
    "rmat-21": {"vertices": 2 ** 21, "edges": 2 ** 25},
    "rmat-22": {"vertices": 2 ** 22, "edges": 2 ** 26},
    "rmat-24": {"vertices": 2 ** 24, "edges": 2 ** 28},
    "rmat-27": {"vertices": 2 ** 27, "edges": 2 ** 31},
    "rmat-32-orig": {"vertices": 2 ** 32, "edges": 2 ** 36},
    "rmat-32": {"vertices": 2 ** 32, "edges": 10 ** 12},
    "rmat-33": {"vertices": 2 ** 33, "edges": 2 ** 40},
}
INPUT_FILE = {
    "twitter-small": "twitter-small/twitter_rv_small-rewritten-new.bin",
    "twitter-full": "twitter-full/twitter_rv-rewritten.bin",
    "friendster": "friendster/friendster-rewritten.bin",
    "livejournal": "livejournal/livejournal.bin",
    "orkut": "orkut/orkut.bin",
    "uk2007": "uk2007-05/uk2007.bin",
}
GRAPHS = ["livejournal", "orkut", "twitter-small"]
TRAVERSALS = ["hilbert", "row-first", "column-first"]
DEFAULT_TRAVERSAL = "row-first"
TILE_DISTRIBUTION_STRATEGIES = ["static", "atomics", "tile-distributor",
                                "hierarchical-tile-distributor"]
DEFAULT_TILE_DISTRIBUTION_STRATEGY = "hierarchical-tile-distributor"
HIERARCHICAL_MAX_META_TILES_PER_DIMENSION = 64
DEFAULT_INSERTION_BATCH_SIZE = 2 ** 20
DEFAULT_TILE_BATCH_SIZE = 256
VERTICES_PER_TILE = 65536
DEFAULT_INTERACTIVE_ALGO = 1
DEFAULT_DYNAMIC_COMPACTION = 0
DEFAULT_ALGORITHM_REEXECUTION = 0
DEFAULT_ENABLE_EDGE_APIS = 1
DEFAULT_REWRITE_IDS = 0
DEFAULT_WRITE_RESULTS = 0
DEFAULT_IN_MEMORY_INGESTION = 0
DEFAULT_WRITE_EDGES = 0
DEFAULT_READ_EDGES = 0
DEFAULT_DELTA = 0.01
DELTAS = [0.00, 0.0025, 0.005, 0.01, 0.02, 0.04, 0.08, 0.16, 0.32, 0.64]
DEFAULT_ALGORITHM_START_VERTEX = 1
MIN_COUNT_EDGES = 0


def data_root():
    return os.environ.get("CYTOM_DATA_ROOT", "/data/graph/")


def out_root():
    return os.path.abspath(os.environ.get("CYTOM_ROOT", os.getcwd()))


def paths():
    r = out_root()
    return {"perf": os.path.join(r, "perf_events"), "edges_out": os.path.join(r, "edges-output"),
            "edges_in": os.path.join(r, "edges-input"),
            "results": os.path.join(r, "results-storage")}


def previous_power_of_2(x):
    """utils.py:73-79."""
    if x == 0 or x == 1:
        return 1
    return 2 ** math.ceil(math.log2(x) - 1)


# ---------------------------------------------------------------- configurations
FIELDS = ("run", "algorithm", "graph", "insertion_batch", "tile_batch", "traversal",
          "dynamic_compaction", "tile_distribution_strategy", "in_memory_ingestion",
          "interactive_algo", "write_edges", "edge_apis", "reexecution", "read_edges",
          "write_results", "deletions", "selective_scheduling", "delta")


def _doubling(spec):
    """'a..b' (doubling from a to b), 'a,b,c' or 'a' (run.py:255-297)."""
    spec = str(spec)
    if "," in spec and ".." not in spec:
        return [int(x.strip()) for x in spec.split(",")]
    lo, hi = (spec.split("..") + [None])[:2] if ".." in spec else (spec, spec)
    lo, hi = int(lo), int(hi)
    out = []
    i = lo
    while i <= hi:
        out.append(i)
        i *= 2
    return out


def config_lists(args, store):
    """ConfigLists.generateConfigs (run.py:245-376): one list per field."""
    L = OrderedDict()
    L["run"] = list(range(int(args.runs)))
    if args.disable_algorithm:
        L["algorithm"] = ["bfs"]
    elif args.algorithm is not None:
        L["algorithm"] = [x.strip() for x in args.algorithm.split(",")]
    else:
        L["algorithm"] = sorted(store.algorithms)
    L["graph"] = ([x.strip() for x in args.graph.split(",")] if args.graph is not None
                  else list(GRAPHS))
    L["insertion_batch"] = (_doubling(args.insertion_batch_size)
                            if args.insertion_batch_size is not None
                            else [DEFAULT_INSERTION_BATCH_SIZE])
    L["tile_batch"] = (_doubling(args.tile_batch_size) if args.tile_batch_size is not None
                       else [DEFAULT_TILE_BATCH_SIZE])
    L["traversal"] = TRAVERSALS if args.enable_traversals else [DEFAULT_TRAVERSAL]
    L["dynamic_compaction"] = (["0", "1"] if args.dynamic_compaction is not None
                               else [DEFAULT_DYNAMIC_COMPACTION])
    L["tile_distribution_strategy"] = (TILE_DISTRIBUTION_STRATEGIES
                                       if args.tile_distribution_strategy is not None
                                       else [DEFAULT_TILE_DISTRIBUTION_STRATEGY])
    L["in_memory_ingestion"] = ["1"] if args.in_memory_ingestion else [DEFAULT_IN_MEMORY_INGESTION]
    if args.disable_algorithm:
        L["interactive_algo"] = ["0"]
    elif args.interactive_algo is not None:
        L["interactive_algo"] = ["0", "1"]
    else:
        L["interactive_algo"] = [DEFAULT_INTERACTIVE_ALGO]
    L["write_edges"] = ["1"] if args.enable_write_edges else [DEFAULT_WRITE_EDGES]
    L["edge_apis"] = ["0", "1"] if args.enable_edge_apis else [DEFAULT_ENABLE_EDGE_APIS]
    L["reexecution"] = (["1"] if args.enable_algorithm_reexecution
                        else [DEFAULT_ALGORITHM_REEXECUTION])
    L["read_edges"] = [args.read_edges] if args.read_edges is not None else [DEFAULT_READ_EDGES]
    L["write_results"] = ["1"] if args.enable_writing_results else [DEFAULT_WRITE_RESULTS]
    L["deletions"] = ["0.01", "0.05", "0.10"] if args.enable_deletions else ["0.00"]
    if args.override_selective_scheduling:
        L["selective_scheduling"] = ["0", "1"]
    elif args.disable_selective_scheduling:
        L["selective_scheduling"] = ["0"]
    else:
        L["selective_scheduling"] = ["1"]
    L["delta"] = DELTAS if args.enable_delta else [DEFAULT_DELTA]
    return L


def configurations(args, store):
    """GraphConfigGenerator.genConfigs (run.py:198-243): the nested loops in
    field order, the run index outermost."""
    L = config_lists(args, store)
    out = [{}]
    for f in FIELDS:
        out = [dict(c, **{f: v}) for c in out for v in L[f]]
    return out


def job_values(args, c):
    """SingleGraphConfig.generateArguments (run.py:56-148) for one configuration."""
    g = c["graph"]
    if g in GRAPH_SETTINGS_DELIM:
        max_vertices = GRAPH_SETTINGS_DELIM[g]["vertices"]
        count_edges = GRAPH_SETTINGS_DELIM[g]["edges"]
        generator = "binary"
        input_file = os.path.join(data_root(), "input", INPUT_FILE[g])
    elif g in GRAPH_SETTINGS_RMAT:
        max_vertices = GRAPH_SETTINGS_RMAT[g]["vertices"]
        count_edges = GRAPH_SETTINGS_RMAT[g]["edges"]
        generator = "rmat"
        input_file = "~"
    else:
        raise ValueError("unknown graph '%s'" % g)
    if args.max_edges > 0:
        count_edges = min(count_edges, args.max_edges)
    meta = 0
    strategy = c["tile_distribution_strategy"]
    if strategy == "hierarchical-tile-distributor":
        strategy = "tile-distributor"
        count_tiles = max_vertices / VERTICES_PER_TILE
        meta = min(previous_power_of_2(count_tiles / 4), HIERARCHICAL_MAX_META_TILES_PER_DIMENSION)
        meta = max(meta, 1)
    managers, inserters = (24, 10) if args.disable_algorithm else (6, 6)
    algo_start = args.algorithm_threshold * count_edges
    if args.algorithm_threshold == -1.0:
        algo_start = MIN_COUNT_EDGES
    algo_start = int(round(algo_start))
    p = paths()
    return OrderedDict([
        ("max-vertices", max_vertices),
        ("vertices-per-tile", VERTICES_PER_TILE),
        ("batch-insertion-count", c["insertion_batch"]),
        ("meta-tile-manager-count", managers),
        ("count-edges", count_edges),
        ("edge-inserters-count", inserters),
        ("use-dynamic-compaction", c["dynamic_compaction"]),
        ("count-parallel-compaction", 48),
        ("enable-edge-apis", c["edge_apis"]),
        ("count-algorithm-executors", 24),
        ("algorithm-tile-batching", c["tile_batch"]),
        ("algorithm-iterations", args.max_iterations),
        ("path-perf-events", p["perf"]),
        ("enable-perf-event-collection", "1" if args.collect_perf_events else "0"),
        ("enable-static-algorithm-before-compaction", "0"),
        ("enable-static-algorithm-after-compaction", "0"),
        ("enable-rewrite-ids", DEFAULT_REWRITE_IDS),
        ("traversal", c["traversal"]),
        ("generator", generator),
        ("binary-input-file", input_file),
        ("algorithm", c["algorithm"]),
        ("deletion-percentage", c["deletions"]),
        ("enable-interactive-algorithm", c["interactive_algo"]),
        ("enable-lazy-algorithm", "0"),
        ("count-algorithm-appliers", 24),
        ("enable-adaptive-scheduling", c["selective_scheduling"]),
        ("tile-distribution-strategy", strategy),
        ("path-to-edges-output", p["edges_out"]),
        ("enable-writing-edges", c["write_edges"]),
        ("enable-in-memory-ingestion", c["in_memory_ingestion"]),
        ("td-count-meta-tiles-per-dimension", meta),
        ("algo-start-id", args.algorithm_start_vertex),
        ("enable-reading-edges", c["read_edges"]),
        ("path-to-edges-input", p["edges_in"]),
        ("enable-algorithm-reexecution", c["reexecution"]),
        ("path-to-results", p["results"]),
        ("enable-write-results", c["write_results"]),
        ("delta", c["delta"]),
        ("start-algorithm-count", algo_start),
        ("output-detailed-tile-stats", "1" if args.detailed_tile_stats else "0"),
    ])


def job_command(job_lists):
    """Command line of one process running the given jobs."""
    return task_command(job_lists)


def tasks(args, store):
    """(algorithms, graph, [job argv lists]) per process of the chosen mode."""
    confs = configurations(args, store)
    jobs = [(c, job_argv(job_values(args, c))) for c in confs]
    if args.concurrent:
        per_graph = OrderedDict()
        for c, j in jobs:
            per_graph.setdefault(c["graph"], []).append((c, j))
        return [([c["algorithm"] for c, _ in v], g, [j for _, j in v])
                for g, v in per_graph.items()]
    return [([c["algorithm"]], c["graph"], [j]) for c, j in jobs]


def build_parser():
    """The flags of run.py:941-997."""
    ap = argparse.ArgumentParser(prog="cytom_amd.launcher")
    for f in ("--clean", "--distclean", "--print-only", "--analyze", "--enable-edge-apis",
              "--enable-algorithm-reexecution", "--enable-writing-results", "--disable-algorithm",
              "--enable-write-edges", "--enable-deletions", "--enable-delta",
              "--override-selective-scheduling", "--wait-after-execution",
              "--in-memory-ingestion", "--enable-traversals", "--debug", "--iostat",
              "--collect-perf-events", "--perf-stat", "--detailed-tile-stats",
              "--disable-selective-scheduling", "--algorithm-store", "--add", "--delete",
              "--clear", "--list", "--sequential", "--concurrent", "--profile"):
        ap.add_argument(f, action="store_true", default=False)
    for f in ("--prefix", "--algorithm", "--graph", "--insertion-batch-size", "--tile-batch-size",
              "--tile-distribution-strategy", "--interactive-algo", "--read-edges",
              "--dynamic-compaction", "--override-algorithm", "--override-graph"):
        ap.add_argument(f, default=None)
    ap.add_argument("--max-edges", default=0, type=int)
    ap.add_argument("--algorithm-threshold", default=0.0, type=float)
    ap.add_argument("--runs", default=1)
    ap.add_argument("--max-iterations", default=100, type=int)
    ap.add_argument("--plot-config", default="")
    ap.add_argument("--algorithm-start-vertex", dest="algorithm_start_vertex",
                    default=DEFAULT_ALGORITHM_START_VERTEX, type=int)
    ap.add_argument("--name", nargs="*")
    ap.add_argument("--path", nargs="*")
    return ap


def main(argv=None, out=print):
    args = build_parser().parse_args(argv)
    store = AlgorithmStore(out=out)
    if args.algorithm_store:
        store.connect(args)
        return 0
    if not (args.sequential or args.profile or args.concurrent or args.print_only):
        out("nothing to do: give --sequential, --concurrent, --profile, --print-only or "
            "--algorithm-store")
        return 0
    todo = tasks(args, store)
    if args.print_only:
        for _, _, jl in todo:
            out(" ".join(shlex.quote(x) for x in job_command(jl)))
        return 0
    algos = []
    for c in configurations(args, store):
        job = job_values(args, c)
        algos.append(Algo(c["algorithm"], c["graph"], job_argv(job), job))
    base = os.path.join(out_root(), "logs", "%s-%d" % (time.strftime("%Y%m%d-%H%M%S"),
                                                       os.getpid()))
    run_logs, k = base, 1
    while os.path.exists(run_logs):  # one directory per launcher run
        run_logs, k = "%s-%d" % (base, k), k + 1
    sched = Scheduler(out=out, log_root=run_logs)
    t0 = time.time()
    failed = sched.run_all(algos) if args.concurrent else sched.profile_all(algos)
    out("----------Overall status--------------")
    out("Tasks: %d, failed: %d, running time (s): %.1f"
        % (len(sched.finished), failed, time.time() - t0))
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
