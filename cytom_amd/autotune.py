"""Geometry and launch-shape autotuner.

Sweeps the tile width (vertices per tile 4096 / 8192 / 16384), the work-chunk
size (chunk_slots 256 / 512 / 1024, i.e. 4 / 8 / 16 64-slot rows per wave
step) and the source bands (1 / 4) on the RMAT update stream of a given
scale and edge factor: 64 batches of E/64 updates (90 % inserts from the
counter-based RMAT generator, 10 % deletions of live edges), BFS, CC, SSSP
and PageRank-delta after every batch, as the bench streams it.  The score
of a shape is the device time (update-batch and run events) of the last 16
batches; the fastest shape per (scale, edge factor) goes into the table that
cytom_create consults when vertices_per_tile is 0.

  python -m cytom_amd.autotune --scales 16,18,20,22 --out tune.json   (GPU)
  python -m cytom_amd.autotune --install tune.json                      (writes
      cytom_amd/tuning_table.json and cytom_amd/csrc/tuning_table.h; rebuild)
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_JSON = os.path.join(HERE, "tuning_table.json")
TABLE_H = os.path.join(HERE, "csrc", "tuning_table.h")

VPTS = (4096, 8192, 16384)
CHUNKS = (256, 512, 1024)
BANDS = (1, 4)
PROGRAMS = ("bfs", "cc", "sssp", "pagerank-delta")


def valid_vpt(scale, vpt):
    # at most 2^26 tiles, as cytom_create keeps for table rows
    V = 1 << scale
    T = max(1, V // vpt)
    return vpt <= V and T * T <= (1 << 26)


def measure(scale, edge_factor, vpt, chunk, bands, batches=64, timed=16, seed=2):
    import cytom_amd as C

    V = 1 << scale
    E = edge_factor << scale
    per = E // batches
    n_del = per // 10
    n_ins = per - n_del
    g = C.Graph(V, vertices_per_tile=vpt, weighted=True, max_iterations=100, delta=1e-7,
                edge_capacity=3 * E)
    g.set_tuning(chunk_slots=chunk, source_bands=bands)
    # BFS/SSSP from the highest-out-degree vertex of the insert stream, as the
    # bench does: the first batch's is close enough for a timing comparison.
    st = C.Stream(seed)
    ids = None
    dev_ms = 0.0
    te = 0
    for b in range(batches):
        first, didx = st.next(n_ins, n_del)
        s, d, w = C.rmat_generate(scale, seed, first, n_ins, weights=True)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(scale, seed, di, len(didx), weights=True)
        if ids is None:
            src = int(np.argmax(np.bincount(s.to_host(), minlength=V)))
            ids = [g.register(p, src) for p in PROGRAMS]
        bs, rs = g.step(ids, s, d, w, n_ins, ds, dd, dw, len(didx))
        if b >= batches - timed:
            dev_ms += bs.t_ingest_ms + sum(r.t_compute_ms for r in rs)
            te += sum(r.traversed_edges for r in rs)
    g.close()
    return dev_ms, te


def sweep(scales, edge_factor, log=print):
    rows = []
    for scale in scales:
        results = []
        for vpt, chunk, bands in itertools.product(VPTS, CHUNKS, BANDS):
            if not valid_vpt(scale, vpt):
                continue
            t0 = time.perf_counter()
            ms, te = measure(scale, edge_factor, vpt, chunk, bands)
            results.append({"vpt": vpt, "chunk_slots": chunk, "source_bands": bands,
                            "device_ms": round(ms, 3), "gteps": round(te / ms / 1e6, 2)})
            log("scale %d vpt %5d chunk %4d bands %d: %8.2f ms  %7.2f GTEPS  (%.1fs)"
                % (scale, vpt, chunk, bands, ms, te / ms / 1e6, time.perf_counter() - t0))
            sys.stdout.flush()
        best = min(results, key=lambda r: r["device_ms"])
        rows.append({"scale": scale, "edge_factor": edge_factor, "vpt": best["vpt"],
                     "chunk_slots": best["chunk_slots"], "source_bands": best["source_bands"],
                     "gteps": best["gteps"], "candidates": results})
    return rows


def install(path, table_json=TABLE_JSON, table_h=TABLE_H):
    rows = json.load(open(path))
    rows = sorted(rows, key=lambda r: (r["edge_factor"], r["scale"]))
    with open(table_json, "w") as f:
        json.dump(rows, f, indent=1)
        f.write("\n")
    lines = [
        "// Launch geometry chosen by cytom_amd/autotune.py per (scale, edge factor)",
        "// of the RMAT update stream.  When vertices_per_tile is 0, cytom_create takes",
        "// the rows of the edge factor nearest the configuration's (edge_capacity / V,",
        "// or 16 without a capacity) and among them the row of the largest scale not",
        "// above log2(V), or the smallest scale when every row is above it.  Generated",
        "// from cytom_amd/tuning_table.json by `python -m cytom_amd.autotune --install`.",
        "#pragma once",
        "",
        "#include <cstdint>",
        "",
        "namespace cytom {",
        "",
        "struct TuneRow {",
        "  uint32_t log2v, edge_factor, vpt, chunk_slots, source_bands;",
        "};",
        "",
        "constexpr TuneRow TUNE_TABLE[] = {",
    ]
    for r in rows:
        lines.append("    {%d, %d, %d, %d, %d},  // %.1f GTEPS" % (
            r["scale"], r["edge_factor"], r["vpt"], r["chunk_slots"], r["source_bands"],
            r["gteps"]))
    lines += ["};", "", "}  // namespace cytom", ""]
    with open(table_h, "w") as f:
        f.write("\n".join(lines))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scales", default="16,18,20,22")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--out", default=None, help="write the sweep's rows as JSON")
    ap.add_argument("--install", default=None,
                    help="install a sweep's JSON as the table (no GPU needed)")
    a = ap.parse_args(argv)
    if a.install:
        for r in install(a.install):
            print("scale %d edge factor %d: vpt %d chunk_slots %d source_bands %d"
                  % (r["scale"], r["edge_factor"], r["vpt"], r["chunk_slots"], r["source_bands"]))
        return 0
    rows = sweep([int(x) for x in a.scales.split(",")], a.edge_factor)
    text = json.dumps(rows, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    else:
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
