"""The C-ABI from C and C++: include/cytom.h compiles as C99 and C++17, the
reference-side binding (tests/shim/execute_shim.cpp, the execute body of a
maintainer's shim) compiles and links against the in-tree library, and the
program statics restate the reference's algorithm classes.  The GPU test
runs the compiled shim on a small binary edge file and checks its results
file against the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import cytom_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "shim"))
import build_shim  # noqa: E402

SHIM = os.path.join(ROOT, "tests", "shim", "execute_shim")
HEADER = os.path.join(ROOT, "include", "cytom.h")


@pytest.mark.parametrize("lang,std", [("c", "c99"), ("c++", "c++17")])
def test_header_compiles(lang, std, tmp_path):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "cytom.h"\nint main(void) { cytom_program_desc d; '
                   'return cytom_program_statics(CYTOM_BFS, &d); }\n')
    cc = "gcc" if lang == "c" else "g++"
    r = subprocess.run([cc, "-std=" + std, "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-fsyntax-only", "-I", os.path.dirname(HEADER), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_builds_and_checks_its_arguments(tmp_path):
    out = build_shim.build(str(tmp_path / "execute_shim"))
    r = subprocess.run([out], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_statics_restate_the_reference_algorithms():
    # examples/bfs.cc:20-32, cc.cc:20-32 (uint32_t, UINT32_MAX), sssp.cc:20-32
    # (float, FLT_MAX, USE_CHANGED false), pagerank-delta.cc:30-65 (three
    # floats, all zero, and the GLOBAL_FIRST_ITERATION rule); every one pulls
    # from active sources only and is synchronous.
    base = C.STATICS_GIVEN | C.NEEDS_SOURCES_ACTIVE
    want = {
        "bfs": (base | C.USE_CHANGED, 4, 1, [2.0 ** 32 - 1, 0, 0], C.FIRST_ITERATION_NONE),
        "cc": (base | C.USE_CHANGED, 4, 1, [2.0 ** 32 - 1, 0, 0], C.FIRST_ITERATION_NONE),
        "sssp": (base, 4, 1, [float(np.finfo(np.float32).max), 0, 0], C.FIRST_ITERATION_NONE),
        "pagerank-delta": (base | C.USE_CHANGED, 12, 3, [0, 0, 0], C.FIRST_ITERATION_PRD_ALL),
        "ppr": (base | C.USE_CHANGED, 12, 3, [0, 0, 0], C.FIRST_ITERATION_NONE),
    }
    for name, (flags, vb, nf, neutral, first) in want.items():
        s = C.program_statics(name)
        assert s == {"flags": flags, "vertex_bytes": vb, "first_iteration": first,
                     "n_fields": nf, "neutral_element": neutral}, name
    with pytest.raises(C.CytomError):
        C.program_statics(7)


def _oracle_results(kind, V, batches, weighted, delta):
    import oracle as O
    og = O.Graph(V, weighted=weighted, max_iterations=100, delta=delta)
    pid = og.register({"bfs": O.BFS, "cc": O.CC, "sssp": O.SSSP, "pagerank-delta": O.PRD}[kind], 0)
    for s, d, w in batches:
        og.update(s, d, w if weighted else None)
        og.execute(pid)
    return og.read(pid)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bfs", "cc", "sssp", "pagerank-delta"])
def test_compiled_shim_matches_the_oracle(kind, tmp_path):
    import oracle as O
    assert C.device_count() > 0, "no HIP device visible"
    assert os.path.exists(SHIM), "build first: tests/shim/build_shim.py"
    scale, seed, n, batch = 12, 11, 60000, 15000
    V = 1 << scale
    s, d = O.rmat_edges(scale, seed, 0, n)
    w = (np.arange(n, dtype=np.uint32) % 63) + 1
    weighted = kind == "sssp"
    if weighted:
        rec = np.zeros(n, np.dtype([("src", "<u4"), ("tgt", "<u4"), ("w", "<f4")]))
        rec["w"] = w.astype(np.float32)
    else:
        rec = np.zeros(n, np.dtype([("src", "<i4"), ("tgt", "<i4")]))
    rec["src"], rec["tgt"] = s, d
    edges = tmp_path / "edges.bin"
    rec.tofile(edges)
    out = tmp_path / "results.bin"
    r = subprocess.run([SHIM, str(edges), str(rec.dtype.itemsize), kind, str(V), str(batch), "0",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    batches = [(s[k:k + batch], d[k:k + batch], w[k:k + batch]) for k in range(0, n, batch)]
    ov = _oracle_results(kind, V, batches, weighted, 0.01)
    if kind == "pagerank-delta":
        got = np.fromfile(out, np.dtype([("delta_sum", "<f4"), ("rank", "<f4"), ("delta", "<f4")]))
        a = got["rank"].astype(np.float64)
        b = ov.astype(np.float64)
        assert np.abs(a / a.sum() - b / b.sum()).sum() <= 1e-6
    elif kind == "sssp":
        got = np.fromfile(out, np.float32)
        want = np.where(ov == 0xFFFFFFFF, np.finfo(np.float32).max, ov.astype(np.float32))
        assert np.array_equal(got, want.astype(np.float32))
    else:
        assert np.array_equal(np.fromfile(out, np.uint32), ov)


@pytest.mark.gpu
def test_registration_checks_stated_statics():
    assert C.device_count() > 0, "no HIP device visible"
    g = C.Graph(1 << 10, weighted=True)
    for name in ("bfs", "cc", "sssp", "pagerank-delta", "ppr"):
        s = C.program_statics(name)
        g.register(name, 0, statics=s)
        # USE_CHANGED does not change the values: either setting registers
        t = dict(s, flags=s["flags"] ^ C.USE_CHANGED)
        g.register(name, 0, statics=t)
        bad = [dict(s, neutral_element=[1.0, 0, 0]),
               dict(s, flags=s["flags"] | C.NEEDS_TARGETS_ACTIVE),
               dict(s, flags=s["flags"] | C.ASYNC),
               dict(s, flags=s["flags"] & ~C.NEEDS_SOURCES_ACTIVE),
               dict(s, vertex_bytes=8),
               dict(s, first_iteration=1 - s["first_iteration"])]
        for b in bad:
            with pytest.raises(C.CytomError) as e:
                g.register(name, 0, statics=b)
            assert e.value.code == -22, (name, b)
