#!/usr/bin/env python3
"""Streaming-graph benchmark on one MI355X.

Workload: the RMAT-22 update stream (4,194,304 vertices, edge factor 16,
a=0.57 b=0.19 c=0.19 d=0.05) cut into 64 batches of 1,048,576 updates
(943,719 inserts from the RMAT stream + 104,857 deletions of uniformly drawn
live edges).  After every batch the engine runs BFS and SSSP (u32 weights in
[1,255]) from the highest out-degree vertex, directed min-label CC and FP32
PageRank-delta (damping 0.85, threshold 1e-7, 100-iteration cap).

One step = one update batch: ingest + all four programs to convergence.
The first --warmup batches are untimed; the next --steps batches are timed.
All inputs are generated into HBM before the timed region.

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


def measured_traffic(args, kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes over the default workload, or None
    when the workload, the kernel or the HIP sources differ from the
    measured ones."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    rec = json.load(open(TRAFFIC_FILE))
    defaults = parse_args([])
    same = all(getattr(args, k) == getattr(defaults, k)
               for k in ("scale", "edge_factor", "batches", "programs", "steps", "warmup",
                         "vertices_per_tile", "chunk_slots", "job_chunks", "blocks_per_cu",
                         "source_bands", "resort_pct", "seed"))
    if not same or rec.get("kernel") != kernel or rec.get("source_digest") != source_digest():
        return None
    return rec["hbm_bytes_per_launch"]


def parse(argv=None):
    args = parse_args(sys.argv[1:] if argv is None else argv)
    if args.config == "rmat27":
        # Preset: every knob the 2^31-update stream needs, unless given.
        given = set(a.split("=")[0] for a in (sys.argv[1:] if argv is None else argv))
        preset = {"--scale": ("scale", 27), "--batches": ("batches", 128),
                  "--programs": ("programs", "cc,pagerank-delta"),
                  "--warmup": ("warmup", 120), "--steps": ("steps", 8),
                  "--warmup-programs": ("warmup_programs", "cc"),
                  "--cpu-baseline": ("cpu_baseline", "off"), "--queries": ("queries", "off"),
                  "--ppr-grad": ("ppr_grad", "off"),
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
    ap.add_argument("--warmup-programs", default="all",
                    help="programs run after each warmup batch (default: all); programs "
                         "re-initialised every batch lose nothing when skipped")
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
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-baseline", choices=["auto", "off"], default="auto")
    ap.add_argument("--cpu-batches", type=int, default=2,
                    help="timed batches of the CPU sample (the last ones of the stream)")
    ap.add_argument("--cpu-threads", type=int, default=0)
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
    pids = {}
    for n in names:
        pids[n] = g.register(n, source if n in ("bfs", "sssp") else 0)
    warm = min(args.warmup, args.batches)
    steps = min(args.steps, args.batches - warm)

    warm_names = names if args.warmup_programs == "all" else \
        [p.strip() for p in args.warmup_programs.split(",") if p.strip()]

    def step(b, acc, run=names):
        s, d, w, ds, dd, dw, nd = batches[b]
        bs = g.update_device(s, d, w, n_ins, ds, dd, dw, nd)
        acc["ingest_ms"] += bs.t_ingest_ms
        acc["deleted"] += bs.deleted
        for n, p in pids.items():
            if n not in run:
                continue
            st = g.execute(p)
            acc["te"][n] += st.traversed_edges
            acc["ms"][n] += st.t_compute_ms
            acc["iters"][n] += st.iterations
            acc["scanned"][n] += st.slots_scanned
        acc["live"] = bs.live_edges

    def new_acc():
        return {"ingest_ms": 0.0, "deleted": 0, "live": 0, "te": {n: 0 for n in names},
                "ms": {n: 0.0 for n in names}, "iters": {n: 0 for n in names},
                "scanned": {n: 0 for n in names}}

    wacc = new_acc()
    for b in range(warm):
        step(b, wacc, warm_names)
        if (b + 1) % 16 == 0:
            log("bench: warmup batch %d/%d" % (b + 1, warm))
    log("bench: warmup %d batches done" % warm)
    g.set_profiling(True)
    acc = new_acc()
    C.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(warm + k, acc)
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
    detail = {
        "gteps_by_program": {n: (acc["te"][n] / (acc["ms"][n] * 1e-3) / 1e9
                                 if acc["ms"][n] else 0.0) for n in names},
        "traversed_edges_by_program": acc["te"],
        "iterations_by_program": acc["iters"],
        "slots_scanned_by_program": acc["scanned"],
        "compute_ms_by_program": {n: round(acc["ms"][n], 3) for n in names},
        "ingest_ms_total": round(acc["ingest_ms"], 3),
        "update_batches_per_s": steps / wall if wall > 0 else 0.0,
        "live_edges_final": acc["live"],
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
        "alg_bytes_per_launch": round(alg / launches) if launches else None,
    }
    res = {
        "value": te_total / wall / 1e9 if wall > 0 else 0.0,
        "ms_per_step": wall * 1e3 / steps if steps else 0.0,
        "steps": steps,
        "warmup": warm,
        "source": source,
        "n_ins": n_ins,
        "n_del": n_del,
        "vpt": g.vpt,
        "tuning": g.get_tuning(),
        "roofline": roofline,
        "detail": detail,
    }
    if args.queries == "on":
        res["detail"]["batched_queries"] = queries_run(args, g, V, source)
    g.close()
    return res


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
    grad = ppr_grad_run(args) if args.ppr_grad == "on" else None
    tq = sum(v["queries"] for v in out.values())
    tb = sum(v["batched_ms"] for v in out.values())
    tp = sum(v["plain_ms"] for v in out.values())
    out["total"] = {"queries": tq, "batched_ms": round(tb, 2), "plain_ms": round(tp, 2),
                    "batched_qps": round(tq / tb * 1e3, 1) if tb else 0.0,
                    "plain_qps": round(tq / tp * 1e3, 1) if tp else 0.0,
                    "speedup": round(tp / tb, 2) if tb else 0.0}
    if grad:
        out["ppr_grad"] = grad
    return out


def ppr_grad_run(args):
    """Differentiable weighted PPR on RMAT-20 (2^24 edges, U(0,1] weights,
    256 teleport vertices, loss = sum of r_i^2 over 1,024 targets): time of
    the operator build (edge upload, CSR/CSC rows, transition probabilities),
    then of one forward and one adjoint on it."""
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
    ta, sweeps_a = _timed(lambda: G.adjoint_device(r_dev, g_dev, dw, ds, tol=1e-10))
    G.close()
    log("bench: PPR gradient on RMAT-20: build %.1f ms, forward %.1f ms (%d sweeps), adjoint %.1f ms "
        "(%d sweeps)" % (tb * 1e3, tf * 1e3, sweeps_f, ta * 1e3, sweeps_a))
    return {"edges": E, "build_ms": round(tb * 1e3, 2), "forward_ms": round(tf * 1e3, 2),
            "forward_sweeps": sweeps_f, "adjoint_ms": round(ta * 1e3, 2),
            "adjoint_sweeps": sweeps_a}


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

    def plain():
        labels = []
        for gi, (scale, ef) in enumerate(specs):
            G = C.Graph(1 << scale, edge_capacity=int((ef << scale) * 1.3) + 1024)
            p = G.register("cc")
            for b in range(nb):
                s, d, n = dev[gi][b]
                G.update_device(s, d, None, n)
                G.execute(p)
            labels.append(G.read(p) if gi < 4 else None)
            G.close()
        return labels

    def union():
        G = C.Graph(V_union, edge_capacity=int(e_total * 1.3) + 1024)
        p = G.register("cc")
        for b in range(nb):
            s, d, n = host[b]
            G.update_device(s, d, None, n)
            G.execute(p)
        lab = G.read(p)
        G.close()
        return lab

    tp, lp = _timed(plain)
    tb, lu = _timed(union)
    same = True
    for gi in range(min(4, len(specs))):
        part = lu[bases[gi]:bases[gi] + (1 << specs[gi][0])].astype(np.int64) - bases[gi]
        same = same and bool(np.array_equal(part, lp[gi].astype(np.int64)))
    log("bench: %d graph tenants (%d edges): union %.1f ms, plain loop %.1f ms"
        % (len(specs), e_total, tb * 1e3, tp * 1e3))
    return {"queries": len(specs), "edges": e_total, "batched_ms": round(tb * 1e3, 2),
            "plain_ms": round(tp * 1e3, 2), "speedup": round(tp / tb, 2),
            "batched_qps": round(len(specs) / tb, 2), "plain_qps": round(len(specs) / tp, 2),
            "sample_equal": same}


def cpu_baseline(args, source):
    """The oracle (multi-threaded C++ restatement) on the host cores: the
    last --cpu-batches batches of the same stream, ingest + all programs."""
    import oracle as O

    E, per_batch, n_ins, n_del = stream_plan(args)
    V = 1 << args.scale
    names = [p.strip() for p in args.programs.split(",") if p.strip()]
    weighted = "sssp" in names
    threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    t_setup = time.perf_counter()
    st = O.Stream(args.seed)
    plan = [st.next(n_ins, n_del) for _ in range(args.batches)]

    def gen(b):
        first, didx = plan[b]
        return O.batch_edges(args.scale, args.seed, first, n_ins, didx, weighted)

    g = O.Graph(V, vertices_per_tile=65536, weighted=weighted, max_iterations=100, delta=1e-7,
                threads=threads)
    first_timed = max(0, args.batches - args.cpu_batches)
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
    cpu_model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": te / wall / 1e9 if wall > 0 else 0.0,
        "unit": "GTEPS",
        "cores": threads,
        "kind": "port",
        "sample": "last %d of %d batches of the same RMAT-%d stream (earlier batches ingested "
                  "untimed, programs from init as after a re-execution), ingest + %s; "
                  "%.1f s timed; %s" % (len(timed_inputs), args.batches, args.scale,
                                        "/".join(names), wall, cpu_model),
    }


def main():
    args = parse()
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
        cpu = cpu_baseline(args, res["source"])
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
                        "%s after every batch" % (args.scale, args.batches, per_batch, n_ins,
                                                  n_del, ", ".join(names)),
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
