#!/usr/bin/env python3
"""Streaming-graph benchmark on one MI355X.

Workload: the RMAT-22 update stream (4,194,304 vertices, edge factor 16,
a=0.57 b=0.19 c=0.19 d=0.05) cut into 64 batches of 1,048,576 updates
(943,719 inserts from the RMAT stream + 104,857 deletions of uniformly drawn
live edges).  After every batch the engine runs BFS and SSSP (u32 weights in
[1,255]) from the highest out-degree vertex, directed min-label CC and FP32
PageRank-delta (damping 0.85, threshold 1e-7, 100-iteration cap).

One step = one update batch: ingest + all four programs to convergence.
Every run streams all --batches batches; the last --steps batches are timed
and the --warmup batches before them are run untimed (the batches before
those are run untimed too, with --prefix-programs).  All inputs are generated
into HBM before the timed region.

value = traversed edges of all programs over the timed batches / wall time
(GTEPS); a traversed edge is a live out-edge of a vertex active in an
iteration, counted by the engine.

After the stream (outside the timed region) the query workload runs once as
batches and once as a plain loop of single-program runs on the same GPU:
256 BFS + 256 SSSP sources and 64 personalized-PageRank teleport sets on the
final graph, and 128 small graph tenants (scale U{10..20}, edge factor
U{8..32}, 8 insert batches, CC after each); detail.batched_queries reports
queries/s of both and the speedup.
"""
import argparse
import concurrent.futures
import glob
import hashlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS per update batch, RMAT-22 incremental BFS/CC/PR; batched vs plain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def source_digest():
    """sha256 over the HIP sources and the C header, in file-name order."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "cytom_amd", "csrc", "*")))
    files.append(os.path.join(ROOT, "include", "cytom.h"))
    for f in files:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gather_prd_rmat22_traffic.json")


# The settings that define the streamed workload (the timed window does not:
# every run streams all batches).
WORKLOAD_KEYS = ("scale", "edge_factor", "batches", "delete_fraction", "programs",
                 "vertices_per_tile", "chunk_slots", "job_chunks", "blocks_per_cu",
                 "source_bands", "resort_pct", "seed")


def measured_traffic(args, kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes over the same workload, or None
    when the workload, the kernel or the HIP sources differ from the
    measured ones."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    rec = json.load(open(TRAFFIC_FILE))
    same = rec.get("workload_key") == {k: getattr(args, k) for k in WORKLOAD_KEYS}
    if not same or rec.get("kernel") != kernel or rec.get("source_digest") != source_digest():
        return None
    return rec["hbm_bytes_per_launch"]


def traffic_summary(args):
    """Writes the traffic record of the PageRank-delta gather from two
    rocprofv3 counter runs of this script (FETCH_SIZE, then WRITE_SIZE, each
    with --calibrate-copy-mib 256): per-launch means over the last
    --traffic-launches dispatches of the gather, each counter scaled by what
    it reads for the calibration copy (a device-to-device copy of known
    size).  The record is keyed on this invocation's workload settings."""
    import csv

    def load(d, counter):
        path = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)[0]
        gather, copy = [], []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
            if "k_gather<3>" in r["Kernel_Name"]:
                gather.append((int(r["Dispatch_Id"]), v))
            elif "copyBuffer" in r["Kernel_Name"]:
                copy.append(v)
        gather.sort()
        last = [v for _, v in gather[-args.traffic_launches:]]
        return sum(last) / len(last), max(copy), len(last)

    copy_kib = args.calibrate_copy_mib * 1024.0
    f, fc, n = load(args.traffic_summary[0], "FETCH_SIZE")
    w, wc, _ = load(args.traffic_summary[1], "WRITE_SIZE")
    fs, ws = copy_kib / fc, copy_kib / wc
    rec = {
        "kernel": "k_gather<pagerank-delta>",
        "workload_key": {k: getattr(args, k) for k in WORKLOAD_KEYS},
        "source_digest": source_digest(),
        "launches": n,
        "fetch_kib_per_launch": round(f, 1),
        "write_kib_per_launch": round(w, 1),
        "fetch_scale": fs,
        "write_scale": ws,
        "hbm_bytes_per_launch": round((f * fs + w * ws) * 1024),
        "commands": [
            "rocprofv3 --pmc %s --kernel-include-regex 'k_gather|copyBuffer' -d gpurun_out/traffic/%s"
            " -o run --output-format csv -- python3 bench.py --cpu-baseline off --queries off"
            " --ppr-grad off --host-paths off --calibrate-copy-mib %d"
            % (c, c, args.calibrate_copy_mib)
            for c in ("FETCH_SIZE", "WRITE_SIZE")] + [
            "python3 bench.py --traffic-summary gpurun_out/traffic/FETCH_SIZE "
            "gpurun_out/traffic/WRITE_SIZE --traffic-launches %d --calibrate-copy-mib %d"
            % (args.traffic_launches, args.calibrate_copy_mib)],
    }
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(rec, fh, indent=1)
        fh.write("\n")
    print(json.dumps(rec, indent=1))
    return 0


def parse(argv=None):
    args = parse_args(sys.argv[1:] if argv is None else argv)
    if args.config == "rmat27":
        # Preset: every knob the 2^31-update stream needs, unless given.
        given = set(a.split("=")[0] for a in (sys.argv[1:] if argv is None else argv))
        preset = {"--scale": ("scale", 27), "--batches": ("batches", 128),
                  "--programs": ("programs", "cc,pagerank-delta"),
                  "--warmup": ("warmup", 2), "--steps": ("steps", 8),
                  "--prefix-programs": ("prefix_programs", "cc"),
                  "--cpu-baseline": ("cpu_baseline", "off"), "--queries": ("queries", "off"),
                  "--ppr-grad": ("ppr_grad", "off"),
                  "--host-paths": ("host_paths", "off"),
                  # 67M tiles of ~29 edges: small chunks put more tiles in
                  # each row-table pass (136.9 vs 131.9 GTEPS at 512 slots)
                  "--chunk-slots": ("chunk_slots", 64),
                  "--delete-fraction": ("delete_fraction", 838860 / 16777216)}
        for flag, (attr, val) in preset.items():
            if flag not in given:
                setattr(args, attr, val)
    return args


def parse_args(argv):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--batches", type=int, default=64)
    ap.add_argument("--delete-fraction", type=float, default=0.1)
    ap.add_argument("--vertices-per-tile", type=int, default=0)
    ap.add_argument("--chunk-slots", type=int, default=0)
    ap.add_argument("--job-chunks", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--source-bands", type=int, default=0)
    ap.add_argument("--resort-pct", type=int, default=0,
                    help="sort-compaction threshold in percent (-1 disables)")
    ap.add_argument("--programs", default="bfs,cc,sssp,pagerank-delta")
    ap.add_argument("--prefix-programs", default="all",
                    help="programs run after each batch before the warmup batches (default: "
                         "all); programs re-initialised every batch lose nothing when skipped")
    ap.add_argument("--config", choices=["rmat22", "rmat27"], default="rmat22",
                    help="rmat27: the RMAT-27 stream (2^31 updates in 128 batches of 2^24, "
                         "CC + PageRank-delta), last 8 batches timed")
    ap.add_argument("--queries", choices=["on", "off"], default="on",
                    help="after the stream: batched vs plain query workload")
    ap.add_argument("--q-bfs", type=int, default=256)
    ap.add_argument("--q-sssp", type=int, default=256)
    ap.add_argument("--q-ppr", type=int, default=64)
    ap.add_argument("--q-graphs", type=int, default=128)
    ap.add_argument("--ppr-grad", choices=["on", "off"], default="on",
                    help="time the differentiable PPR forward/adjoint on RMAT-20")
    ap.add_argument("--api", choices=["step", "calls"], default="step",
                    help="step: one cytom_step per batch (batch and first iteration groups "
                         "enqueued together); calls: update_device then execute per program")
    ap.add_argument("--concurrent-runs", choices=["on", "off"], default="on",
                    help="on: a step runs its programs concurrently, one stream each; off: one "
                         "after the other on the context's stream")
    ap.add_argument("--profile-every", type=int, default=4,
                    help="time every n-th gather launch of every run with HIP events, the "
                         "phase rotating per run (1: every launch; each timed launch costs a "
                         "few microseconds)")
    ap.add_argument("--host-paths", choices=["on", "off"], default="on",
                    help="also stream an RMAT-(scale-2) graph through both host paths "
                         "(step vs separate calls) after the timed stream")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-baseline", choices=["auto", "off"], default="auto")
    ap.add_argument("--cpu-batches", type=int, default=2,
                    help="timed batches of the CPU sample (the last ones of the stream, "
                         "inside the GPU's timed window)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--traffic-summary", nargs=2, default=None, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/gather_prd_rmat22_traffic.json from two rocprofv3 "
                         "counter runs (no GPU work)")
    ap.add_argument("--traffic-launches", type=int, default=4026,
                    help="gather dispatches averaged by --traffic-summary (the timed ones)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (the reference "
                         "kernel for HBM traffic counters)")
    return ap.parse_args(argv)


def calibration_copy(mib):
    import cytom_amd as C

    n = mib << 20
    a = C.DeviceArray(n, np.uint8)
    b = C.DeviceArray(n, np.uint8)
    rc = C.lib().cytom_memcpy_d2d(b.ptr, a.ptr, n)
    if rc:
        raise C.CytomError(rc, "calibration copy failed")
    C.synchronize()
    log("bench: calibration copy of %d MiB done" % mib)


def stream_plan(args):
    E = args.edge_factor << args.scale
    per_batch = E // args.batches
    n_del = int(per_batch * args.delete_fraction)
    n_ins = per_batch - n_del
    return E, per_batch, n_ins, n_del


def gpu_run(args):
    import cytom_amd as C

    E, per_batch, n_ins, n_del = stream_plan(args)
    V = 1 << args.scale
    names = [p.strip() for p in args.programs.split(",") if p.strip()]
    weighted = "sssp" in names
    t_setup = time.perf_counter()
    # Inputs: inserts from the counter-based generator, deletions sampled from
    # the live set on the host and materialised on the device.
    stream = C.Stream(args.seed)
    batches = []
    need_source = any(n in ("bfs", "sssp") for n in names)
    outdeg = np.zeros(V, np.int64) if need_source else None
    for b in range(args.batches):
        first, didx = stream.next(n_ins, n_del)
        s, d, w = C.rmat_generate(args.scale, args.seed, first, n_ins, weights=weighted)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(args.scale, args.seed, di, len(didx), weights=weighted)
        batches.append((s, d, w, ds, dd, dw, len(didx)))
        if need_source:
            outdeg += np.bincount(s.to_host(), minlength=V)
        if (b + 1) % 32 == 0:
            log("bench: generated %d/%d batches" % (b + 1, args.batches))
    # BFS/SSSP source: highest out-degree vertex of the insert stream, ties to
    # the lowest id.
    source = int(np.argmax(outdeg)) if need_source else 0
    del outdeg
    log("bench: generated %d batches (%d inserts + %d deletions each) in %.1fs; source=%d"
        % (args.batches, n_ins, n_del, time.perf_counter() - t_setup, source))
    g = C.Graph(V, vertices_per_tile=args.vertices_per_tile, weighted=weighted,
                max_iterations=100, delta=1e-7, edge_capacity=int(3 * E))
    g.set_tuning(args.chunk_slots, args.job_chunks, args.blocks_per_cu, args.source_bands,
                 C.RESORT_OFF if args.resort_pct < 0 else args.resort_pct)
    g.set_concurrent_runs(args.concurrent_runs == "on")
    pids = {}
    for n in names:
        pids[n] = g.register(n, source if n in ("bfs", "sssp") else 0)
    steps = max(0, min(args.steps, args.batches))
    warm = max(0, min(args.warmup, args.batches - steps))
    prefix = args.batches - steps - warm
    prefix_names = names if args.prefix_programs == "all" else \
        [p.strip() for p in args.prefix_programs.split(",") if p.strip()]

    def step(b, acc, run=names):
        s, d, w, ds, dd, dw, nd = batches[b]
        run_names = [n for n in pids if n in run]
        if args.api == "step":
            bs, rs = g.step([pids[n] for n in run_names], s, d, w, n_ins, ds, dd, dw, nd)
            stats = dict(zip(run_names, rs))
            syncs = bs.host_syncs
        else:
            bs = g.update_device(s, d, w, n_ins, ds, dd, dw, nd)
            stats = {n: g.execute(pids[n]) for n in run_names}
            syncs = bs.host_syncs + sum(1 for _ in run_names)  # at least one per run
        acc["ingest_ms"] += bs.t_ingest_ms
        acc["deleted"] += bs.deleted
        acc["syncs"] += syncs
        te = 0
        for n in run_names:
            st = stats[n]
            acc["te"][n] += st.traversed_edges
            acc["ms"][n] += st.t_compute_ms
            acc["iters"][n] += st.iterations
            acc["scanned"][n] += st.slots_scanned
            te += st.traversed_edges
        acc["live"] = bs.live_edges
        return te

    def new_acc():
        return {"ingest_ms": 0.0, "deleted": 0, "live": 0, "syncs": 0,
                "te": {n: 0 for n in names},
                "ms": {n: 0.0 for n in names}, "iters": {n: 0 for n in names},
                "scanned": {n: 0 for n in names}}

    pacc = new_acc()
    for b in range(prefix):
        step(b, pacc, prefix_names)
        if (b + 1) % 16 == 0:
            log("bench: untimed batch %d/%d" % (b + 1, prefix))
    wacc = new_acc()
    for b in range(prefix, prefix + warm):
        step(b, wacc)
    log("bench: %d untimed + %d warmup batches done" % (prefix, warm))
    g.set_profiling(True, every=args.profile_every)
    acc = new_acc()
    per_batch_te, per_batch_s = [], []
    C.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        tb0 = time.perf_counter()
        per_batch_te.append(step(prefix + warm + k, acc))
        per_batch_s.append(time.perf_counter() - tb0)  # execute() ends synchronised
        if (k + 1) % 8 == 0:
            log("bench: timed batch %d/%d" % (k + 1, steps))
    C.synchronize()
    wall = time.perf_counter() - t0
    te_total = sum(acc["te"].values())
    kinds = {n: C.PROGRAMS[n] for n in names}
    prof = {n: g.kernel_profile(kinds[n]) for n in names}
    dom = max(names, key=lambda n: prof[n][0])
    ms, launches, alg = prof[dom]
    achieved = alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    cpu_n = max(0, min(args.cpu_batches, steps))
    detail = {
        "batches_streamed": args.batches,
        "timed_batches": [prefix + warm + 1, args.batches] if steps else [],
        "gteps_by_program": {n: (acc["te"][n] / (acc["ms"][n] * 1e-3) / 1e9
                                 if acc["ms"][n] else 0.0) for n in names},
        "traversed_edges_by_program": acc["te"],
        "iterations_by_program": acc["iters"],
        "slots_scanned_by_program": acc["scanned"],
        "slots_per_traversed_edge": {n: round(acc["scanned"][n] / acc["te"][n], 3)
                                     if acc["te"][n] else None for n in names},
        "compute_ms_by_program": {n: round(acc["ms"][n], 3) for n in names},
        "ingest_ms_total": round(acc["ingest_ms"], 3),
        "api": args.api,
        "host_syncs_per_batch": round(acc["syncs"] / steps, 2) if steps and args.api == "step"
        else None,
        # Wall time not covered by the batch's and the runs' device events:
        # host work and launch gaps (work lists are built between them).
        "host_overhead_ms_per_batch": round(
            (wall * 1e3 - acc["ingest_ms"] - sum(acc["ms"].values())) / steps, 3)
        if steps else None,
        "update_batches_per_s": steps / wall if wall > 0 else 0.0,
        "live_edges_final": acc["live"],
        # The GPU rate over exactly the batches the CPU baseline times.
        "gteps_last_batches": {
            "batches": [args.batches - cpu_n + 1, args.batches] if cpu_n else [],
            "value": (sum(per_batch_te[steps - cpu_n:]) / sum(per_batch_s[steps - cpu_n:]) / 1e9
                      if cpu_n and sum(per_batch_s[steps - cpu_n:]) > 0 else None)},
        "gather_kernel_by_program": {
            n: {"ms": round(prof[n][0], 3), "launches": prof[n][1],
                "alg_gbytes": round(prof[n][2] / 1e9, 3),
                "achieved_gbs": round(prof[n][2] / (prof[n][0] * 1e-3) / 1e9, 1)
                if prof[n][0] else 0.0}
            for n in names},
    }
    roofline = {
        "bound": "hbm",
        "kernel": "k_gather<%s>" % dom,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(args, "k_gather<%s>" % dom),
        "avg_launch_us": round(ms * 1e3 / launches, 2) if launches else None,
        "timed_launches": launches,
        "timed_every": args.profile_every,
        "alg_bytes_per_launch": round(alg / launches) if launches else None,
    }
    # Final values of every program, for the comparison against the CPU
    # oracle on the same stream (done in the CPU leg).
    final = {n: g.read(p) for n, p in pids.items()}
    res = {
        "value": te_total / wall / 1e9 if wall > 0 else 0.0,
        "ms_per_step": wall * 1e3 / steps if steps else 0.0,
        "steps": steps,
        "warmup": warm,
        "prefix": prefix,
        "source": source,
        "n_ins": n_ins,
        "n_del": n_del,
        "vpt": g.vpt,
        "tuning": g.get_tuning(),
        "roofline": roofline,
        "detail": detail,
        "final": final,
    }
    res["grad_arrays"] = None
    if args.queries == "on":
        res["detail"]["batched_queries"], res["grad_arrays"] = queries_run(args, g, V, source)
    g.close()
    if args.host_paths == "on":
        res["detail"]["host_paths"] = api_comparison(args, scale=max(args.scale - 2, 10))
    return res


def api_comparison(args, scale=20, batches=16):
    """The same RMAT stream through both host paths (a graph each): per batch,
    the wall time, the device time of the batch and the runs (HIP events),
    their difference, and the host synchronisations of the step path."""
    import cytom_amd as C

    V = 1 << scale
    E = args.edge_factor << scale
    n_del = int(E // batches * args.delete_fraction)
    n_ins = E // batches - n_del
    names = [p.strip() for p in args.programs.split(",") if p.strip()]
    weighted = "sssp" in names
    stream = C.Stream(args.seed + 1)
    data = []
    for b in range(batches):
        first, didx = stream.next(n_ins, n_del)
        s, d, w = C.rmat_generate(scale, args.seed + 1, first, n_ins, weights=weighted)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(scale, args.seed + 1, di, len(didx), weights=weighted)
        data.append((s, d, w, ds, dd, dw, len(didx)))
    out = {"workload": "RMAT-%d, %d batches x %d updates, %s; batches 5-%d timed"
                       % (scale, batches, n_ins + n_del, ", ".join(names), batches)}
    finals = {}
    # Step graphs replay while the argument pointers repeat: the graph path
    # stages each batch into fixed device buffers (untimed, as any resident
    # input would be).
    staged = [C.DeviceArray(max(n_ins, n_del, 1), np.uint32) for _ in range(6)]
    for api in ("calls", "step", "step_graphs"):
        g = C.Graph(V, vertices_per_tile=args.vertices_per_tile, weighted=weighted,
                    max_iterations=100, delta=1e-7, edge_capacity=int(3 * E))
        ids = [g.register(n, 0) for n in names]
        if api == "step_graphs":
            g.set_step_graphs(True)
        wall = dev = 0.0
        syncs = 0
        for b, (s, d, w, ds, dd, dw, nd) in enumerate(data):
            if api == "step_graphs":
                for buf, a in zip(staged, (s, d, w, ds, dd, dw)):
                    if a is not None:
                        buf.copy_from(a)
                s, d, w, ds, dd, dw = [buf if a is not None else None
                                       for buf, a in zip(staged, (s, d, w, ds, dd, dw))]
            C.synchronize()
            t0 = time.perf_counter()
            if api in ("step", "step_graphs"):
                bs, rs = g.step(ids, s, d, w, n_ins, ds, dd, dw, nd)
            else:
                bs = g.update_device(s, d, w, n_ins, ds, dd, dw, nd)
                rs = [g.execute(p) for p in ids]
            t = time.perf_counter() - t0
            if b >= 4:
                wall += t * 1e3
                dev += bs.t_ingest_ms + sum(r.t_compute_ms for r in rs)
                syncs += bs.host_syncs
        finals[api] = [g.read(p) for p in ids]
        n = batches - 4
        out[api] = {"wall_ms_per_batch": round(wall / n, 3), "device_ms_per_batch": round(dev / n, 3),
                    "host_overhead_ms_per_batch": round((wall - dev) / n, 3)}
        if api != "calls":
            out[api]["host_syncs_per_batch"] = round(syncs / n, 2)
        if api == "step_graphs":
            cap, rep, _ = g.step_graph_stats()
            out[api]["graphs_recorded"] = cap
            out[api]["graphs_replayed"] = rep
        g.close()
    out["values_equal"] = all(np.array_equal(a, b) and np.array_equal(a, c) for a, b, c in
                              zip(finals["calls"], finals["step"], finals["step_graphs"]))
    out["host_overhead_saved_ms_per_batch"] = round(
        out["calls"]["host_overhead_ms_per_batch"] - out["step"]["host_overhead_ms_per_batch"], 3)
    out["step_graphs_saved_ms_per_batch"] = round(
        out["step"]["wall_ms_per_batch"] - out["step_graphs"]["wall_ms_per_batch"], 3)
    return out


def _timed(fn):
    import cytom_amd as C

    C.synchronize()
    t0 = time.perf_counter()
    out = fn()
    C.synchronize()
    return time.perf_counter() - t0, out


def queries_run(args, g, V, source):
    """Query workload on the graph left by the stream: BFS and SSSP from
    sources uniform among vertices with in+out degree > 0 and personalized
    PageRank from teleport sets of U{1..1024} such vertices, each run once as
    one batch (cytom_batched_execute) and once as a plain per-query loop of
    single-program runs on the same GPU; plus the small-graph tenants."""
    import cytom_amd as C

    rng = np.random.default_rng(args.seed + 100)
    deg_o, deg_i = g.degrees()
    live = np.nonzero(deg_o.astype(np.int64) + deg_i)[0].astype(np.uint32)
    n_sssp = args.q_sssp if g.weighted else 0
    sets = {
        "bfs": [("bfs", int(s)) for s in rng.choice(live, args.q_bfs)],
        "sssp": [("sssp", int(s)) for s in rng.choice(live, n_sssp)],
        "ppr": [("ppr", rng.choice(live, int(rng.integers(1, 1025))))
                for _ in range(args.q_ppr)],
    }
    out = {}
    pids = {}
    for kind, qs in sets.items():
        if not qs:
            continue
        tb, st = _timed(lambda: g.batched_execute(qs))
        sample = [0, len(qs) // 2, len(qs) - 1]
        got = {i: g.read_query(i) for i in sample}
        if kind not in pids:
            pids[kind] = g.register(kind, int(live[0]))
        pid = pids[kind]

        def plain():
            te = 0
            for _, src in qs:
                g.set_sources(pid, src)
                te += g.execute(pid).traversed_edges
            return te

        tp, te_plain = _timed(plain)
        same = True
        for i in sample:
            g.set_sources(pid, qs[i][1])
            g.execute(pid)
            same = same and bool(np.array_equal(got[i], g.read(pid)))
        out[kind] = {"queries": len(qs), "batched_ms": round(tb * 1e3, 2),
                     "plain_ms": round(tp * 1e3, 2), "speedup": round(tp / tb, 2),
                     "batched_qps": round(len(qs) / tb, 1), "plain_qps": round(len(qs) / tp, 1),
                     "batched_traversed_edges": st.traversed_edges,
                     "plain_traversed_edges": te_plain, "sample_equal": same}
        log("bench: %d %s queries: batched %.1f ms, plain loop %.1f ms" % (len(qs), kind,
                                                                          tb * 1e3, tp * 1e3))
    if args.q_graphs:
        out["graphs"] = tenants_run(args)
    grad, grad_arrays = ppr_grad_run(args) if args.ppr_grad == "on" else (None, None)
    tq = sum(v["queries"] for v in out.values())
    tb = sum(v["batched_ms"] for v in out.values())
    tp = sum(v["plain_ms"] for v in out.values())
    out["total"] = {"queries": tq, "batched_ms": round(tb, 2), "plain_ms": round(tp, 2),
                    "batched_qps": round(tq / tb * 1e3, 1) if tb else 0.0,
                    "plain_qps": round(tq / tp * 1e3, 1) if tp else 0.0,
                    "speedup": round(tp / tb, 2) if tb else 0.0}
    if grad:
        out["ppr_grad"] = grad
    return out, grad_arrays


def ppr_grad_run(args):
    """Differentiable weighted PPR on RMAT-20 (2^24 edges, U(0,1] weights,
    256 teleport vertices, loss = sum of r_i^2 over 1,024 targets): time of
    the operator build (edge upload, CSR/CSC rows, transition probabilities),
    then of one forward and one adjoint on it, and central differences
    (h = 1e-3) on 512 edges.  Returns the record and the arrays the CPU leg
    checks against the FP64 restatement."""
    import cytom_amd as C
    from cytom_amd import autograd as A

    scale, V, E = 20, 1 << 20, 1 << 24
    s, d, _ = C.rmat_generate(scale, args.seed + 300, 0, E)
    rng = np.random.default_rng(args.seed + 300)
    w = (1.0 - rng.random(E)).astype(np.float32)
    t = np.zeros(V, np.float32)
    t[rng.choice(V, 256, replace=False)] = 1.0 / 256
    targets = rng.integers(0, V, 1024)
    sh, dh = s.to_host(), d.to_host()
    tb, G = _timed(lambda: A.PPRGraph(V, sh, dh, w))
    t_dev = C.DeviceArray.from_host(t)
    r_dev = C.DeviceArray(V, np.float32)
    tf, sweeps_f = _timed(lambda: G.forward_device(t_dev, r_dev, tol=1e-10))
    r = r_dev.to_host()
    g = np.zeros(V, np.float32)
    np.add.at(g, targets, 2.0 * r[targets])
    g_dev = C.DeviceArray.from_host(g)
    dw = C.DeviceArray(E, np.float32)
    ds = C.DeviceArray(V, np.float32)
    # lambda is of the order of the ranks (~1e-6): 1e-14 keeps its relative
    # error near 1e-8
    ta, sweeps_a = _timed(lambda: G.adjoint_device(r_dev, g_dev, dw, ds, tol=1e-14))
    # Central differences of the loss on 512 edges drawn with probability
    # proportional to |dL/dw| (most edges' gradients are below the FP32
    # resolution of the loss), against the adjoint's gradient; a fixed sweep
    # count, so both sides of a difference run the same iteration.
    grad = dw.to_host()
    n_fd, sweeps_fd, h = 512, 160, 1e-3
    pw = np.abs(grad).astype(np.float64)
    picks = rng.choice(E, n_fd, replace=False, p=pw / pw.sum())

    def loss(wv):
        G.set_weights(wv)
        G.forward_device(t_dev, r_dev, tol=0.0, max_iter=sweeps_fd)
        rv = r_dev.to_host().astype(np.float64)
        return float(np.sum(rv[targets] ** 2))

    t_fd = time.perf_counter()
    fd = []
    for e in picks:
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd.append((loss(wp) - loss(wm)) / (2 * h))
    t_fd = time.perf_counter() - t_fd
    fd = np.asarray(fd)
    rel_fd = float(np.linalg.norm(grad[picks] - fd) / max(np.linalg.norm(fd), 1e-300))
    G.close()
    log("bench: PPR gradient on RMAT-20: build %.1f ms, forward %.1f ms (%d sweeps), adjoint %.1f ms "
        "(%d sweeps); central differences on %d edges: rel. error %.2e"
        % (tb * 1e3, tf * 1e3, sweeps_f, ta * 1e3, sweeps_a, n_fd, rel_fd))
    rec = {"edges": E, "build_ms": round(tb * 1e3, 2), "forward_ms": round(tf * 1e3, 2),
           "forward_sweeps": sweeps_f, "adjoint_ms": round(ta * 1e3, 2),
           "adjoint_sweeps": sweeps_a,
           "fd_check": {"edges": n_fd, "h": h, "sweeps": sweeps_fd,
                        "rel_error_vs_central_differences": rel_fd,
                        "seconds": round(t_fd, 2)}}
    arrays = {"V": V, "src": sh, "dst": dh, "w": w, "t": t, "targets": targets, "r": r,
              "dw": grad, "ds": ds.to_host(), "picks": picks, "fd": fd}
    return rec, arrays


def tenants_run(args):
    """--q-graphs small graphs (scale U{10..20}, edge factor U{8..32}, own
    RMAT seeds), each streamed in 8 equal insert batches with CC after every
    batch: a plain loop of one context per graph against one context holding
    their disjoint union (vertex ranges side by side), which runs every
    tenant's batch b as one update and one CC run."""
    import cytom_amd as C

    rng = np.random.default_rng(args.seed + 200)
    specs = [(int(rng.integers(10, 21)), int(rng.integers(8, 33))) for _ in range(args.q_graphs)]
    nb = 8
    dev = []  # per graph, per batch: (src, dst) device arrays
    host = []  # per batch: union (src, dst) host arrays
    base = 0
    bases = []
    for gi, (scale, ef) in enumerate(specs):
        e_b = (ef << scale) // nb
        seed = args.seed + 1000 + gi
        bb = []
        for b in range(nb):
            s, d, _ = C.rmat_generate(scale, seed, b * e_b, e_b)
            bb.append((s, d, e_b))
        dev.append(bb)
        bases.append(base)
        base += 1 << scale
    V_union = base
    for b in range(nb):
        us = np.concatenate([dev[gi][b][0].to_host() + np.uint32(bases[gi])
                             for gi in range(len(specs))])
        ud = np.concatenate([dev[gi][b][1].to_host() + np.uint32(bases[gi])
                             for gi in range(len(specs))])
        host.append((C.DeviceArray.from_host(us), C.DeviceArray.from_host(ud), len(us)))
    e_total = sum(h[2] for h in host)

    # Contexts are created (and their CC registered) before either timed
    # region, so both legs time only the update batches and CC runs.
    t_ctx = time.perf_counter()
    plain_graphs = []
    for gi, (scale, ef) in enumerate(specs):
        G = C.Graph(1 << scale, edge_capacity=int((ef << scale) * 1.3) + 1024)
        plain_graphs.append((G, G.register("cc")))
    G_union = C.Graph(V_union, edge_capacity=int(e_total * 1.3) + 1024)
    p_union = G_union.register("cc")
    C.synchronize()
    t_ctx = time.perf_counter() - t_ctx

    def plain():
        for gi, (G, p) in enumerate(plain_graphs):
            for b in range(nb):
                s, d, n = dev[gi][b]
                G.update_device(s, d, None, n)
                G.execute(p)
        return [G.read(p) for G, p in plain_graphs[:4]]

    def union():
        for b in range(nb):
            s, d, n = host[b]
            G_union.update_device(s, d, None, n)
            G_union.execute(p_union)
        return G_union.read(p_union)

    tp, lp = _timed(plain)
    tb, lu = _timed(union)
    for G, _ in plain_graphs:
        G.close()
    G_union.close()
    same = True
    for gi in range(min(4, len(specs))):
        part = lu[bases[gi]:bases[gi] + (1 << specs[gi][0])].astype(np.int64) - bases[gi]
        same = same and bool(np.array_equal(part, lp[gi].astype(np.int64)))
    log("bench: %d graph tenants (%d edges): union %.1f ms, plain loop %.1f ms"
        % (len(specs), e_total, tb * 1e3, tp * 1e3))
    return {"queries": len(specs), "edges": e_total, "batched_ms": round(tb * 1e3, 2),
            "plain_ms": round(tp * 1e3, 2), "speedup": round(tp / tb, 2),
            "context_setup_ms_untimed": round(t_ctx * 1e3, 2),
            "batched_qps": round(len(specs) / tb, 2), "plain_qps": round(len(specs) / tp, 2),
            "sample_equal": same}


def cpu_model_name():
    name = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name


def cpu_baseline(args, source, gpu_final):
    """The oracle (multi-threaded C++ restatement) on the host cores: the
    last --cpu-batches batches of the same stream (inside the GPU's timed
    window), ingest + all programs.  Afterwards, outside the CPU timing, the
    GPU's final values are compared with the oracle's: BFS/CC/SSSP exactly,
    PageRank-delta by normalised L1 against the oracle's FP32 run and against
    its FP64 restatement on the same final graph."""
    import oracle as O

    E, per_batch, n_ins, n_del = stream_plan(args)
    V = 1 << args.scale
    names = [p.strip() for p in args.programs.split(",") if p.strip()]
    weighted = "sssp" in names
    threads = args.cpu_threads or O.host_threads()
    t_setup = time.perf_counter()
    st = O.Stream(args.seed)
    plan = [st.next(n_ins, n_del) for _ in range(args.batches)]

    def gen(b):
        first, didx = plan[b]
        return O.batch_edges(args.scale, args.seed, first, n_ins, didx, weighted)

    g = O.Graph(V, vertices_per_tile=65536, weighted=weighted, max_iterations=100, delta=1e-7,
                threads=threads)
    n_timed = max(1, min(args.cpu_batches, args.batches))
    first_timed = args.batches - n_timed
    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        futs = [ex.submit(gen, b) for b in range(args.batches)]
        for b in range(first_timed):
            g.update(*futs[b].result())
            futs[b] = None
        timed_inputs = [futs[b].result() for b in range(first_timed, args.batches)]
    kinds = {"bfs": O.BFS, "cc": O.CC, "sssp": O.SSSP, "pagerank-delta": O.PRD}
    pids = {n: g.register(kinds[n], source if n in ("bfs", "sssp") else 0) for n in names}
    log("bench: cpu baseline setup %.1fs (%d threads)" % (time.perf_counter() - t_setup, threads))
    te = 0
    t0 = time.perf_counter()
    for inp in timed_inputs:
        g.update(*inp)
        for n, p in pids.items():
            te += g.execute(p).traversed_edges
    wall = time.perf_counter() - t0
    nproc = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    out = {
        "value": te / wall / 1e9 if wall > 0 else 0.0,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": "batches %d-%d of the %d-batch RMAT-%d stream (the last %d of the GPU's timed "
                  "window; earlier batches ingested untimed, programs from init as after a "
                  "re-execution), ingest + %s; %.1f s timed; %d threads (nproc %d, affinity %d, "
                  "OMP_NUM_THREADS %s); %s"
                  % (first_timed + 1, args.batches, args.batches, args.scale, n_timed,
                     "/".join(names), wall, threads, nproc, affinity,
                     os.environ.get("OMP_NUM_THREADS", "unset"), cpu_model_name()),
    }
    # Parity of the GPU's final values against the oracle (not timed).
    parity = {}
    for n, p in pids.items():
        if n not in gpu_final:
            continue
        ov, gv = g.read(p), gpu_final[n]
        if n == "pagerank-delta":
            parity["pagerank-delta_l1_vs_oracle_fp32"] = l1n(gv, ov)
        else:
            parity[n + "_equal"] = bool(np.array_equal(ov, gv))
    if "pagerank-delta" in gpu_final:
        t1 = time.perf_counter()
        ls, ld, _ = g.live_edges()
        r64, it64 = O.prd_f64(V, ls, ld, max_iterations=100, delta=1e-7, threads=threads)
        parity["pagerank-delta_l1_vs_fp64"] = l1n(gpu_final["pagerank-delta"], r64)
        parity["oracle_fp32_l1_vs_fp64"] = l1n(g.read(pids["pagerank-delta"]), r64)
        parity["fp64_iterations"] = int(it64)
        parity["pagerank-delta_fixed_point_residual"] = O.prd_residual(
            V, ls, ld, gpu_final["pagerank-delta"], threads=threads)
        log("bench: FP64 PageRank-delta on the final graph %.1fs" % (time.perf_counter() - t1))
    out["final_parity"] = parity
    return out


def grad_fp64_check(a):
    """The CPU leg's check of the PPR-gradient record: the FP64 restatement's
    forward and adjoint on the same RMAT-20 operator, the relative errors of
    the GPU's r, dL/dw and dL/ds against it, and of the FP64 dL/dw against
    the same central differences."""
    from oracle import ppr_grad as PG

    t0 = time.perf_counter()
    V, s, d = a["V"], a["src"], a["dst"]
    w64 = a["w"].astype(np.float64)
    r64 = PG.forward(V, s, d, w64, a["t"].astype(np.float64), tol=1e-13)
    _, g64 = PG.loss_sq(r64, a["targets"])
    dw64, ds64, _ = PG.adjoint(V, s, d, w64, r64, g64, tol=1e-13)

    def rel(x, y):
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64)
        return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))

    out = {"rel_error_r_vs_fp64": rel(a["r"], r64),
           "rel_error_dw_vs_fp64_adjoint": rel(a["dw"], dw64),
           "rel_error_ds_vs_fp64_adjoint": rel(a["ds"], ds64),
           "fp64_adjoint_rel_error_vs_central_differences": rel(dw64[a["picks"]], a["fd"]),
           "seconds": round(time.perf_counter() - t0, 1)}
    log("bench: PPR gradient vs FP64: r %.2e, dL/dw %.2e, dL/ds %.2e (%.1fs)"
        % (out["rel_error_r_vs_fp64"], out["rel_error_dw_vs_fp64_adjoint"],
           out["rel_error_ds_vs_fp64_adjoint"], out["seconds"]))
    return out


def l1n(a, b):
    """L1 distance of two rank vectors after L1 normalisation."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def main():
    args = parse()
    if args.traffic_summary:
        return traffic_summary(args)
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        return 2
    E, per_batch, n_ins, n_del = stream_plan(args)
    if args.calibrate_copy_mib:
        calibration_copy(args.calibrate_copy_mib)
    res = gpu_run(args)
    cpu = None
    if args.cpu_baseline == "auto":
        cpu = cpu_baseline(args, res["source"], res["final"])
        res["detail"]["final_parity"] = cpu.pop("final_parity")
        if res["grad_arrays"] is not None:
            res["detail"]["batched_queries"]["ppr_grad"]["fp64_check"] = \
                grad_fp64_check(res["grad_arrays"])
    names = [p.strip() for p in args.programs.split(",") if p.strip()]
    line = {
        "metric": METRIC,
        "value": round(res["value"], 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": res["steps"],
        "warmup": res["warmup"],
        "ms_per_step": round(res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/f32",
        "data": "synthetic (counter-based RMAT stream, seed %d)" % args.seed,
        "config": {
            "workload": "RMAT-%d stream, %d batches x %d updates (%d inserts + %d deletions); "
                        "%s after every batch; all %d batches streamed, batches %d-%d timed"
                        % (args.scale, args.batches, per_batch, n_ins, n_del, ", ".join(names),
                           args.batches, res["prefix"] + res["warmup"] + 1, args.batches),
            "scale": args.scale,
            "edge_factor": args.edge_factor,
            "vertices_per_tile": res["vpt"],
            "tuning": res["tuning"],
            "source": res["source"],
            "parallelism": "single GPU",
        },
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "detail": res["detail"],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
