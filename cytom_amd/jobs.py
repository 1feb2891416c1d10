"""The reference's job argument format.

One job is 80 arguments: 40 ``--key value`` pairs in the fixed order that
scripts/run.py:107-148 (SingleGraphConfig.generateArguments) writes and
main/execute.h:49-90 (parseOption) declares; a process gets k jobs as
1 + 80 k argv entries (core/jobs.h:31-40, NUM_OF_ARGS = 80).  execute.h reads
the string-valued options by argv position (argv[26], [36], [38], [40], [54],
[56], [68], [72]), so the order is part of the format and parse_job checks it.
The edge source (generator, binary-input-file) of the first job feeds every
job of the process (Jobs::parse_options, jobs.h:168-244).
"""
from collections import OrderedDict

NUM_OF_ARGS = 80

U64, INT, DOUBLE, STR = "u64", "int", "double", "str"

# (key, type) in argv order.
JOB_KEYS = (
    ("max-vertices", U64),
    ("vertices-per-tile", U64),
    ("batch-insertion-count", U64),
    ("meta-tile-manager-count", U64),
    ("count-edges", U64),
    ("edge-inserters-count", U64),
    ("use-dynamic-compaction", INT),
    ("count-parallel-compaction", U64),
    ("enable-edge-apis", INT),
    ("count-algorithm-executors", U64),
    ("algorithm-tile-batching", U64),
    ("algorithm-iterations", U64),
    ("path-perf-events", STR),
    ("enable-perf-event-collection", INT),
    ("enable-static-algorithm-before-compaction", INT),
    ("enable-static-algorithm-after-compaction", INT),
    ("enable-rewrite-ids", INT),
    ("traversal", STR),
    ("generator", STR),
    ("binary-input-file", STR),
    ("algorithm", STR),
    ("deletion-percentage", DOUBLE),
    ("enable-interactive-algorithm", INT),
    ("enable-lazy-algorithm", INT),
    ("count-algorithm-appliers", U64),
    ("enable-adaptive-scheduling", INT),
    ("tile-distribution-strategy", STR),
    ("path-to-edges-output", STR),
    ("enable-writing-edges", INT),
    ("enable-in-memory-ingestion", INT),
    ("td-count-meta-tiles-per-dimension", U64),
    ("algo-start-id", U64),
    ("enable-reading-edges", INT),
    ("path-to-edges-input", STR),
    ("enable-algorithm-reexecution", INT),
    ("path-to-results", STR),
    ("enable-write-results", INT),
    ("delta", DOUBLE),
    ("start-algorithm-count", U64),
    ("output-detailed-tile-stats", INT),
)
assert 2 * len(JOB_KEYS) == NUM_OF_ARGS

TRAVERSALS = ("hilbert", "row-first", "column-first")      # execute.h:191-210
GENERATORS = ("rmat", "binary")                            # execute.h:212-230
TILE_DISTRIBUTIONS = ("tile-distributor", "atomics", "static")  # execute.h:268-287


class JobError(ValueError):
    pass


def _convert(key, typ, text):
    try:
        if typ == U64:
            v = int(text, 10)
            if v < 0 or v >= 1 << 64:
                raise ValueError
            return v
        if typ == INT:
            return int(text, 10)
        if typ == DOUBLE:
            return float(text)
        return text
    except ValueError:
        raise JobError("the argument ('%s') for option '--%s' is invalid" % (text, key))


def job_argv(values):
    """80 argv strings of one job from a key -> value mapping (every key)."""
    out = []
    for key, _ in JOB_KEYS:
        if key not in values:
            raise JobError("the option '--%s' is required but missing" % key)
        out += ["--" + key, str(values[key])]
    return out


def split_jobs(argv):
    """argv without the program name -> one 80-entry list per job."""
    if len(argv) == 0 or len(argv) % NUM_OF_ARGS:
        raise JobError("Wrong number of arguments, %d, expected multiple of : %d"
                       % (len(argv) + 1, NUM_OF_ARGS))
    return [argv[i:i + NUM_OF_ARGS] for i in range(0, len(argv), NUM_OF_ARGS)]


def parse_job(args):
    """80 argv strings -> OrderedDict key -> typed value, validated as
    execute.h:45-362 does (lower-cased enum strings, allowed values)."""
    if len(args) != NUM_OF_ARGS:
        raise JobError("a job has %d arguments, not %d" % (len(args), NUM_OF_ARGS))
    job = OrderedDict()
    for i, (key, typ) in enumerate(JOB_KEYS):
        flag, text = args[2 * i], args[2 * i + 1]
        if flag != "--" + key:
            raise JobError("argument %d is '%s' where '--%s' belongs" % (2 * i + 1, flag, key))
        job[key] = _convert(key, typ, text)
    for key, allowed in (("traversal", TRAVERSALS), ("generator", GENERATORS),
                         ("tile-distribution-strategy", TILE_DISTRIBUTIONS)):
        job[key] = job[key].lower()
        if job[key] not in allowed:
            raise JobError("Wrong option given for %s: %s" % (key, job[key]))
    return job


def parse_jobs(argv):
    return [parse_job(a) for a in split_jobs(argv)]
