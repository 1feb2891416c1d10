"""Parity of the HIP path against the CPU oracle, called through the C-ABI.

BFS, CC and SSSP values and the store contents must be bit-exact after every
batch; PageRank-delta ranks within 1e-6 L1 of the oracle after L1
normalisation (ranks are unnormalised, sum ~ V).
"""
import collections

import numpy as np
import pytest

import cytom_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

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


# ------------------------------------------------------------------ generator
@pytest.mark.parametrize("scale", [1, 10, 16, 22, 27])
def test_rmat_device_matches_oracle(scale):
    first, n = 123456789 % (1 << 20), 200000
    s, d, w = C.rmat_generate(scale, 2, first, n, weights=True)
    es, ed = O.rmat_edges(scale, 2, first, n)
    assert np.array_equal(s.to_host(), es)
    assert np.array_equal(d.to_host(), ed)
    assert np.array_equal(w.to_host(), O.weights_u32(2, first, n))


def test_rmat_gather_by_index_matches_oracle():
    idx = np.random.default_rng(5).integers(0, 1 << 30, 50000).astype(np.uint64)
    di = C.DeviceArray.from_host(idx)
    s, d, w = C.rmat_gather(22, 2, di, len(idx), weights=True)
    hs, hd = s.to_host(), d.to_host()
    for k in range(0, len(idx), 997):
        es, ed = O.rmat_edges(22, 2, int(idx[k]), 1)
        assert hs[k] == es[0] and hd[k] == ed[0]
        assert w.to_host()[k] == O.weights_u32(2, int(idx[k]), 1)[0]


# ------------------------------------------------------------------ streams
def run_stream_parity(scale, vpt, batches, n_ins, n_del, seed=2, edge_capacity=0,
                      max_iterations=100, tuning=None, delete_first=False):
    V = 1 << scale
    st = O.Stream(seed)
    og = O.Graph(V, vertices_per_tile=min(vpt, 65536), weighted=True,
                 max_iterations=max_iterations)
    gg = C.Graph(V, vertices_per_tile=vpt, weighted=True, edge_capacity=edge_capacity,
                 max_iterations=max_iterations)
    if tuning:
        gg.set_tuning(**tuning)
    progs = [("bfs", O.BFS, C.BFS, 0), ("cc", O.CC, C.CC, 0), ("sssp", O.SSSP, C.SSSP, 0),
             ("prd", O.PRD, C.PRD, 0)]
    oid = {n: og.register(k, s) for n, k, _, s in progs}
    gid = {n: gg.register(k, s) for n, _, k, s in progs}
    for b in range(batches):
        first, didx = st.next(n_ins, n_del if (b or delete_first) else 0)
        s, d, w, ds, dd, dw = O.batch_edges(scale, seed, first, n_ins, didx, True)
        nf = og.update(s, d, w, ds, dd, dw)
        bs = gg.update(s, d, w, ds, dd, dw)
        assert bs.deleted == nf and bs.missing == len(ds) - nf
        os_, od_, ow_ = og.live_edges()
        gs_, gd_, gw_ = gg.export_edges()
        assert multiset(os_, od_, ow_) == multiset(gs_, gd_, gw_), "store multiset"
        oo, oi = og.degrees()
        go, gi = gg.degrees()
        assert np.array_equal(oo, go) and np.array_equal(oi, gi), "degrees"
        for name, _, _, _ in progs:
            ost = og.execute(oid[name])
            gst = gg.execute(gid[name])
            ov = og.read(oid[name])
            gv = gg.read(gid[name])
            if name == "prd":
                err = l1n(gv, ov)
                assert err <= PRD_L1_TOL, (b, err)
            else:
                bad = np.nonzero(ov != gv)[0]
                assert bad.size == 0, (b, name, bad[:10], ov[bad[:10]], gv[bad[:10]])
                assert gst.converged == ost.converged
            assert gst.traversed_edges > 0 or ost.traversed_edges == 0
    return og, gg, oid, gid


def test_stream_parity_small_tiles():
    run_stream_parity(scale=12, vpt=64, batches=6, n_ins=3000, n_del=300)


def test_stream_parity_wide_tiles():
    run_stream_parity(scale=13, vpt=4096, batches=6, n_ins=8000, n_del=800)


def test_stream_parity_deletions_in_first_batch():
    # Deletions before the first run, at a rate that takes the last out-edge
    # of many vertices: the PageRank-delta DELETE hook and re-execution test
    # see the out-degree with the deleted edge still counted.
    run_stream_parity(scale=11, vpt=256, batches=4, n_ins=3000, n_del=1500, delete_first=True)


def test_stream_parity_with_pool_growth():
    # A tiny initial pool forces repacks (global compaction) mid-stream.
    run_stream_parity(scale=11, vpt=256, batches=5, n_ins=4000, n_del=1000, edge_capacity=1024)


@pytest.mark.parametrize("tuning", [
    dict(source_bands=1), dict(source_bands=4), dict(source_bands=16),
    dict(chunk_slots=64, job_chunks=1, source_bands=2),
    dict(chunk_slots=1024, job_chunks=128, blocks_per_cu=1),
    dict(resort_pct=C.RESORT_OFF), dict(resort_pct=1), dict(resort_pct=1, resort_max_slots=700)])
def test_stream_parity_launch_shapes(tuning):
    # Work-list shapes (chunk size, chunks per job, source bands, grid) and
    # the sort-compaction policy (off, after every batch, a small staging
    # budget that defers tiles to later batches) change only the slot order
    # and the order of the exact reductions, never the values.  Tiles of
    # 2048 vertices collect ~500 inserts per batch, past the 256-slot tail
    # that makes a tile eligible for sort-compaction.
    run_stream_parity(scale=13, vpt=2048, batches=5, n_ins=8000, n_del=1600, tuning=tuning)


def test_launch_shapes_agree_at_rmat20():
    # Size-independent property at a large graph: every launch shape gives
    # bit-identical values for all four programs, batch after batch.  The
    # PageRank-delta sums are exact fixed point, so its whole trajectory
    # (iterations, traversed edges) is identical too; the BFS/CC/SSSP insert
    # hooks relax concurrently, so only their fixpoint values are compared.
    scale, V = 20, 1 << 20
    st = O.Stream(2)
    # 64-slot chunks in jobs of 8192 chunks: heavy columns take several
    # 1024-chunk passes per job.
    shapes = [dict(source_bands=1), dict(source_bands=8, resort_pct=C.RESORT_OFF),
              dict(source_bands=16, chunk_slots=256, job_chunks=16, resort_pct=5),
              dict(chunk_slots=64, job_chunks=8192)]
    graphs = []
    for t in shapes:
        g = C.Graph(V, weighted=True)
        g.set_tuning(**t)
        graphs.append((g, [g.register(C.BFS, 0), g.register(C.CC), g.register(C.SSSP, 0),
                           g.register(C.PRD)]))
    for b in range(4):
        first, didx = st.next(1 << 20, (1 << 17) if b else 0)
        s, d, w = C.rmat_generate(scale, 2, first, 1 << 20, weights=True)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(scale, 2, di, len(didx), weights=True)
        vals = []
        for g, ids in graphs:
            g.update_device(s, d, w, 1 << 20, ds, dd, dw, len(didx))
            row = []
            for p in ids:
                rs = g.execute(p)
                row.append((rs.iterations, rs.traversed_edges, g.read(p)))
            vals.append(row)
        for row in vals[1:]:
            for k in range(4):
                assert np.array_equal(row[k][2], vals[0][k][2]), (b, k)
            assert row[3][0] == vals[0][3][0] and row[3][1] == vals[0][3][1], b


def test_insert_only_cc_rmat16():
    # RMAT-16, 16 insert-only batches of 65,536 edges, CC labels bit-exact
    # after every batch; the textbook checker confirms the oracle.
    scale, V = 16, 1 << 16
    og = O.Graph(V, vertices_per_tile=65536)
    gg = C.Graph(V)
    op, gp = og.register(O.CC), gg.register(C.CC)
    for b in range(16):
        s, d = O.rmat_edges(scale, 2, b * 65536, 65536)
        og.update(s, d)
        gg.update(s, d)
        og.execute(op)
        gst = gg.execute(gp)
        ov, gv = og.read(op), gg.read(gp)
        assert np.array_equal(ov, gv), b
        assert gst.converged == 1 and gst.reexecuted == 0
    ls, ld, _ = og.live_edges()
    assert np.array_equal(gv, O.check_cc(V, ls, ld))


def test_medium_stream_against_checkers():
    scale, V = 18, 1 << 18
    st = O.Stream(11)
    gg = C.Graph(V, weighted=True)
    og = O.Graph(V, vertices_per_tile=16384, weighted=True, threads=8)
    ids = [gg.register(C.BFS, 0), gg.register(C.CC), gg.register(C.SSSP, 0),
           gg.register(C.PRD)]
    op = og.register(O.PRD)
    for b in range(4):
        first, didx = st.next(250000, 25000 if b else 0)
        s, d, w, ds, dd, dw = O.batch_edges(scale, 11, first, 250000, didx, True)
        gg.update(s, d, w, ds, dd, dw)
        og.update(s, d, w, ds, dd, dw)
        for p in ids:
            gg.execute(p)
        og.execute(op)
        ls, ld, lw = gg.export_edges()
        assert np.array_equal(gg.read(ids[0]), O.check_bfs(V, ls, ld, 0))
        assert np.array_equal(gg.read(ids[1]), O.check_cc(V, ls, ld))
        assert np.array_equal(gg.read(ids[2]), O.check_sssp(V, ls, ld, lw, 0))
        assert l1n(gg.read(ids[3]), og.read(op)) <= PRD_L1_TOL


# ------------------------------------------------------------------ edge cases
def test_empty_and_missing_and_invalid():
    g = C.Graph(256, vertices_per_tile=64, weighted=True)
    p = g.register(C.SSSP, 3)
    e = np.empty(0, np.uint32)
    st = g.update(e, e, e)
    assert st.inserted == 0 and st.live_edges == 0
    g.execute(p)
    assert g.read(p)[3] == 0 and (np.delete(g.read(p), 3) == 0xFFFFFFFF).all()
    # invalid vertex: rejected, nothing modified
    with pytest.raises(C.CytomError) as ei:
        g.update(np.array([1], np.uint32), np.array([256], np.uint32), np.array([1], np.uint32))
    assert ei.value.code == C.EINVAL
    # a zero weight: refused by the binding, and by the library itself
    with pytest.raises(ValueError):
        g.update(np.array([1], np.uint32), np.array([2], np.uint32), np.array([0], np.uint32))
    one, zero = C.DeviceArray.from_host(np.array([1], np.uint32)), C.DeviceArray.from_host(
        np.array([0], np.uint32))
    with pytest.raises(C.CytomError) as ei:
        g.update_device(one, C.DeviceArray.from_host(np.array([2], np.uint32)), zero, 1)
    assert ei.value.code == C.EINVAL
    assert g.export_edges()[0].size == 0
    # self loops, duplicates, and a deletion of an edge inserted in the same batch
    s = np.array([3, 3, 3, 5, 5, 7], np.uint32)
    d = np.array([3, 5, 5, 7, 7, 9], np.uint32)
    w = np.array([1, 4, 2, 1, 1, 3], np.uint32)
    st = g.update(s, d, w, np.array([5, 100], np.uint32), np.array([7, 101], np.uint32),
                  np.array([1, 1], np.uint32))
    assert st.deleted == 1 and st.missing == 1 and st.live_edges == 5
    g.execute(p)
    v = g.read(p)
    assert v[3] == 0 and v[5] == 2 and v[7] == 3 and v[9] == 6
    # delete both copies of 3->5: 5 becomes unreachable
    st = g.update(e, e, e, np.array([3, 3], np.uint32), np.array([5, 5], np.uint32), None)
    assert st.deleted == 2
    g.execute(p)
    v = g.read(p)
    assert v[5] == 0xFFFFFFFF and v[9] == 0xFFFFFFFF


def test_iteration_cap_leftovers():
    og = O.Graph(64, vertices_per_tile=64, max_iterations=3)
    gg = C.Graph(64, vertices_per_tile=64, max_iterations=3)
    op, gp = og.register(O.BFS, 0), gg.register(C.BFS, 0)
    s = np.arange(8, -1, -1).astype(np.uint32)
    og.update(s, s + 1)
    gg.update(s, s + 1)
    for _ in range(4):
        ost, gst = og.execute(op), gg.execute(gp)
        assert ost.converged == gst.converged and ost.iterations == gst.iterations
        assert np.array_equal(og.read(op), gg.read(gp))
    assert gg.read(gp)[:10].tolist() == list(range(10))


def test_prd_insert_only_hooks_path():
    # Batches without deletions take the edgeChanged(INSERT) path of
    # PageRank-delta instead of re-initialising.
    scale, V = 11, 1 << 11
    og = O.Graph(V, vertices_per_tile=512)
    gg = C.Graph(V, vertices_per_tile=512)
    op, gp = og.register(O.PRD), gg.register(C.PRD)
    for b in range(4):
        s, d = O.rmat_edges(scale, 3, b * 4000, 4000)
        og.update(s, d)
        gg.update(s, d)
        og.execute(op)
        gg.execute(gp)
        assert l1n(gg.read(gp), og.read(op)) <= PRD_L1_TOL, b


def test_prd_fixed_point_fraction_bits():
    # The PageRank-delta neighbour sums are int64 fixed point with
    # 60 - ceil(log2 2V) fraction bits: 43 at RMAT-16, 32 at RMAT-27.  Forcing
    # 32 bits at RMAT-16 reproduces the RMAT-27 quantum (2^-32 per
    # contribution) on a graph the FP64 restatement solves in seconds: the
    # ranks must stay within the 1e-6 bound of the FP64 run and of the FP32
    # oracle.  16 bits (quantum 1.5e-5) shows the check is sensitive to the
    # quantisation: its error is far larger.
    scale, V = 16, 1 << 16
    st = O.Stream(21)
    og = O.Graph(V, vertices_per_tile=16384)
    graphs = {}
    for bits in (0, 32, 16):
        g = C.Graph(V)
        if bits:
            g.set_tuning(prd_frac_bits=bits)
        assert g.get_tuning()["prd_frac_bits"] == (bits or 43)
        graphs[bits] = (g, g.register(C.PRD))
    op = og.register(O.PRD)
    for b in range(3):
        first, didx = st.next(300000, 30000 if b else 0)
        s, d, _, ds, dd, _ = O.batch_edges(scale, 21, first, 300000, didx, False)
        og.update(s, d, None, ds, dd, None)
        for g, p in graphs.values():
            g.update(s, d, None, ds, dd, None)
            g.execute(p)
        og.execute(op)
    ls, ld, _ = og.live_edges()
    r64, _ = O.prd_f64(V, ls, ld)
    err = {bits: (l1n(g.read(p), r64), l1n(g.read(p), og.read(op)))
           for bits, (g, p) in graphs.items()}
    print("PageRank-delta L1 (vs FP64, vs FP32 oracle) by fraction bits:", err, flush=True)
    for bits in (0, 32):
        assert err[bits][0] <= PRD_L1_TOL and err[bits][1] <= PRD_L1_TOL, (bits, err[bits])
    assert err[16][0] > 10 * max(err[32][0], 1e-9)
    with pytest.raises(C.CytomError):
        graphs[0][0].set_tuning(prd_frac_bits=44)  # above 60 - ceil(log2 2V) at V = 2^16


# ------------------------------------------------------- personalized PageRank
def test_ppr_stream_parity():
    # Single-query PPR program vs the oracle after every batch of a stream
    # with deletions (from scratch per batch), several teleport sets.
    scale, V = 13, 1 << 13
    st = O.Stream(4)
    og = O.Graph(V, vertices_per_tile=1024)
    gg = C.Graph(V, vertices_per_tile=1024)
    op, gp = og.register(O.PPR, 3), gg.register(C.PPR, 3)
    sets = [[3], [1, 1, 2, 900, 8000], list(range(5, V, 101))]
    for b in range(4):
        first, didx = st.next(20000, 2000 if b else 0)
        s, d, w, ds, dd, dw = O.batch_edges(scale, 4, first, 20000, didx, False)
        og.update(s, d, None, ds, dd, None)
        gg.update(s, d, None, ds, dd, None)
        for tel in sets:
            og.set_sources(op, tel)
            gg.set_sources(gp, tel)
            ost, gst = og.execute(op), gg.execute(gp)
            assert gst.converged == ost.converged
            err = l1n(gg.read(gp), og.read(op))
            assert err <= PRD_L1_TOL, (b, tel[:2], err)


# ------------------------------------------------------------ batched queries
def _stream_graph(scale, batches, n_ins, n_del, weighted=True, seed=6, **kw):
    V = 1 << scale
    st = O.Stream(seed)
    g = C.Graph(V, weighted=weighted, **kw)
    for b in range(batches):
        first, didx = st.next(n_ins, n_del if b else 0)
        s, d, w, ds, dd, dw = O.batch_edges(scale, seed, first, n_ins, didx, weighted)
        g.update(s, d, w if weighted else None, ds, dd, dw if weighted else None)
    return g


def test_batched_queries_equal_single_runs():
    scale = 14
    V = 1 << scale
    g = _stream_graph(scale, 3, 60000, 6000)
    rng = np.random.default_rng(3)
    deg_o, deg_i = g.degrees()
    live = np.nonzero(deg_o + deg_i)[0]
    bfs_src = rng.choice(live, 70)
    bfs_src[5] = bfs_src[6]  # the same source twice in one group
    bfs_src[7] = np.nonzero(deg_o + deg_i == 0)[0][0]  # an isolated vertex
    sssp_src = rng.choice(live, 20)
    ppr_sets = [rng.choice(live, int(rng.integers(1, 60))) for _ in range(18)]
    ppr_sets[0] = np.array([live[0], live[0], live[1]])
    queries = ([("bfs", s) for s in bfs_src] + [("sssp", s) for s in sssp_src] +
               [("ppr", t) for t in ppr_sets])
    # interleave kinds so grouping by kind is exercised
    order = rng.permutation(len(queries))
    queries = [queries[i] for i in order]
    st = g.batched_execute(queries)
    assert st.traversed_edges > 0
    pid = {k: g.register(k, int(live[0])) for k in ("bfs", "sssp", "ppr")}
    ls, ld, lw = g.export_edges()
    for i, (kind, src) in enumerate(queries):
        got = g.read_query(i)
        g.set_sources(pid[kind], src)
        g.execute(pid[kind])
        ref = g.read(pid[kind])
        assert np.array_equal(got, ref), (i, kind)
        if kind == "bfs" and i % 7 == 0:
            assert np.array_equal(got, O.check_bfs(V, ls, ld, int(src)))
        if kind == "sssp" and i % 5 == 0:
            assert np.array_equal(got, O.check_sssp(V, ls, ld, lw, int(src)))


def test_batched_ppr_iteration_cap_and_errors():
    g = _stream_graph(12, 2, 20000, 2000, max_iterations=3)
    sets = [[1, 2, 3], [7], list(range(0, 4096, 9))]
    g.batched_execute([("ppr", t) for t in sets])
    p = g.register("ppr", 0)
    for i, t in enumerate(sets):
        g.set_sources(p, t)
        rs = g.execute(p)
        assert rs.iterations == 3
        assert np.array_equal(g.read_query(i), g.read(p))
    with pytest.raises(RuntimeError):
        g.batched_execute([("bfs", [1, 2])])
    with pytest.raises(RuntimeError):
        g.batched_execute([("cc", [1])])
    with pytest.raises(RuntimeError):
        g.batched_execute([("bfs", [1 << 20])])
    gu = C.Graph(256, vertices_per_tile=64)
    gu.update(np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    with pytest.raises(RuntimeError):
        gu.batched_execute([("sssp", [0])])
    r = gu.batched_execute([("bfs", [0]), ("ppr", [0])])
    assert list(gu.read_query(0)[:4]) == [0, 1, 2, 0xFFFFFFFF]


def test_batched_queries_with_hub_vertices():
    # In-degree above the hub threshold (2048) takes the block-per-vertex
    # pull path; a dense frontier switches from push to pull.
    scale, V = 12, 1 << 12
    rng = np.random.default_rng(8)
    s, d = O.rmat_edges(scale, 8, 0, 30000)
    hub_in = rng.integers(0, V, 6000).astype(np.uint32)
    hub_out = rng.integers(0, V, 6000).astype(np.uint32)
    s = np.concatenate([s, hub_in, np.full(6000, 7, np.uint32)])
    d = np.concatenate([d, np.full(6000, 7, np.uint32), hub_out])
    w = (1 + rng.integers(0, 255, len(s))).astype(np.uint32)
    g = C.Graph(V, vertices_per_tile=1024, weighted=True)
    g.update(s, d, w)
    srcs = rng.integers(0, V, 40)
    queries = ([("bfs", x) for x in srcs] + [("sssp", x) for x in srcs[:20]] +
               [("ppr", rng.integers(0, V, int(k))) for k in (1, 3, 40, 300)])
    g.batched_execute(queries)
    pid = {k: g.register(k, 0) for k in ("bfs", "sssp", "ppr")}
    for i, (kind, src) in enumerate(queries):
        g.set_sources(pid[kind], src)
        g.execute(pid[kind])
        assert np.array_equal(g.read_query(i), g.read(pid[kind])), (i, kind)


def test_batched_bfs_deep_graph_rounds():
    # A 700-vertex path plus sparse random shortcuts: BFS depths in the
    # hundreds, so the batched BFS runs many rounds of device-driven levels
    # (8, then 4 at a time, alternating frontier lists) and levels past the
    # last one.  Levels equal the textbook BFS; 70 queries make two groups.
    V = 1 << 12
    rng = np.random.default_rng(12)
    path = np.arange(700, dtype=np.uint32)
    s = np.concatenate([path[:-1], rng.integers(0, V, 400).astype(np.uint32)])
    d = np.concatenate([path[1:], rng.integers(700, V, 400).astype(np.uint32)])
    g = C.Graph(V, vertices_per_tile=1024)
    g.update(s, d)
    srcs = np.concatenate([[0, 1, 350, 699], rng.integers(0, V, 66)]).astype(np.uint32)
    g.batched_execute([("bfs", int(x)) for x in srcs])
    for i, x in enumerate(srcs):
        assert np.array_equal(g.read_query(i), O.check_bfs(V, s, d, int(x))), (i, int(x))


def test_batched_queries_large_graph_route():
    # Above 2^24 live edges SSSP groups run on the query tiles, eight queries
    # per tile-engine run, and PPR queries run one per tile-engine run; BFS
    # keeps the bit-parallel groups.  Values still equal the single-query runs.
    scale, V = 20, 1 << 20
    g = C.Graph(V, weighted=True)
    n = 1 << 22
    for b in range(5):
        s, d, w = C.rmat_generate(scale, 12, b * n, n, weights=True)
        g.update_device(s, d, w, n)
    assert g.live_edge_count() > (1 << 24)
    rng = np.random.default_rng(2)
    src = rng.integers(0, 64, 6)  # low ids: high degree under RMAT
    # Eleven SSSP queries: one run of eight and one of three (a repeated
    # source, a high-id vertex that reaches little); five PPR queries (sets
    # with a repeated vertex, a large set, a sink).
    sssp = [int(x) for x in src[:3]] + [int(src[0]), V - 1] + [int(x) for x in
                                                                 rng.integers(0, V, 6)]
    queries = ([("sssp", x) for x in sssp] + [("bfs", x) for x in src[3:]] +
               [("ppr", src), ("ppr", [int(src[0])]), ("ppr", [int(src[1]), int(src[1]), 7]),
                ("ppr", rng.integers(0, V, 300)), ("ppr", [V - 1])])
    g.batched_execute(queries)
    pid = {k: g.register(k, 0) for k in ("bfs", "sssp", "ppr")}
    for i, (kind, sv) in enumerate(queries):
        g.set_sources(pid[kind], sv)
        g.execute(pid[kind])
        assert np.array_equal(g.read_query(i), g.read(pid[kind])), (i, kind)


# ------------------------------------------------ differentiable weighted PPR
def _weighted(V, E, seed, n_tel=16):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, V, E).astype(np.int32)
    d = rng.integers(0, V, E).astype(np.int32)
    w = (1.0 - rng.random(E)).astype(np.float32)  # U(0, 1]
    t = np.zeros(V, np.float32)
    t[rng.choice(V, n_tel, replace=False)] = 1.0 / n_tel
    return s, d, w, t


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_ppr_grad_forward_adjoint_and_autograd():
    import torch
    from cytom_amd import autograd as A
    from oracle import ppr_grad as G
    V, E = 1 << 11, 20000
    s, d, w, t = _weighted(V, E, 1)
    targets = np.random.default_rng(2).integers(0, V, 256)
    r_ref = G.forward(V, s, d, w.astype(np.float64), t.astype(np.float64))
    tw = torch.from_numpy(w).requires_grad_(True)
    tt = torch.from_numpy(t).requires_grad_(True)
    r = A.ppr(torch.from_numpy(s), torch.from_numpy(d), tw, tt)
    assert _rel(r.detach().numpy(), r_ref) < 1e-5
    (r[torch.from_numpy(targets)] ** 2).sum().backward()
    _, g = G.loss_sq(r_ref, targets)
    dw_ref, ds_ref, _ = G.adjoint(V, s, d, w.astype(np.float64), r_ref, g)
    assert _rel(tw.grad.numpy(), dw_ref) < 1e-4
    assert _rel(tt.grad.numpy(), ds_ref) < 1e-4
    # Deterministic: the same inputs give bit-identical results.
    r2 = A.ppr(torch.from_numpy(s), torch.from_numpy(d), tw.detach(), tt.detach())
    assert torch.equal(r2, r.detach())


def test_ppr_weights_must_be_positive_and_finite():
    # A zero, negative, NaN or infinite weight is refused (P = w / W would be
    # NaN or negative); a refused set_weights leaves the operator unchanged; a
    # non-finite teleport vector makes the solve fail instead of reporting a
    # converged NaN result.
    from cytom_amd import autograd as A
    V, E = 1 << 10, 5000
    s, d, w, t = _weighted(V, E, 9)
    for bad in (0.0, -1.0, np.nan, np.inf):
        wb = w.copy()
        wb[17] = bad
        with pytest.raises(C.CytomError) as ei:
            A.PPRGraph(V, s, d, wb)
        assert ei.value.code == C.EINVAL, bad
    G = A.PPRGraph(V, s, d, w)
    r0 = G.forward(t)
    wb = w.copy()
    wb[3] = np.nan
    with pytest.raises(C.CytomError) as ei:
        G.set_weights(wb)
    assert ei.value.code == C.EINVAL and "positive" in str(ei.value)
    assert np.array_equal(G.forward(t), r0)
    tb = t.copy()
    tb[5] = np.nan
    with pytest.raises(C.CytomError) as ei:
        G.forward(tb)
    assert ei.value.code == C.ERANGE
    assert np.array_equal(G.forward(t), r0)
    G.close()


def test_ppr_grad_finite_differences_gpu():
    # Central differences of the FP32 forward (h = 1e-3) on sampled edges.
    from cytom_amd import autograd as A
    V, E = 1 << 12, 40000
    s, d, w, t = _weighted(V, E, 3, n_tel=64)
    targets = np.random.default_rng(4).integers(0, V, 1024)
    G = A.PPRGraph(V, s, d, w)
    r = G.forward(t)
    g = np.zeros(V, np.float32)
    np.add.at(g, targets, 2.0 * r[targets])
    grad, _ = G.adjoint(r, g)

    def L(wv):
        G.set_weights(wv)
        rv = G.forward(t).astype(np.float64)
        return float(np.sum(rv[targets] ** 2))

    h = 1e-3
    picks = np.random.default_rng(5).choice(E, 64, replace=False)
    fd = []
    for e in picks:
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd.append((L(wp) - L(wm)) / (2 * h))
    assert _rel(grad[picks], fd) < 2e-2


def test_ppr_graph_handle_matches_one_shot_calls():
    # The graph handle builds its rows once; forward and adjoint through it
    # must equal the one-shot entry points bit for bit (same rows, same
    # summation order), and set_weights must equal a fresh build with those
    # weights.  An endpoint >= V is rejected before any kernel indexes by it.
    import ctypes
    from cytom_amd import autograd as A
    V, E = 1 << 12, 40000
    s, d, w, t = _weighted(V, E, 6, n_tel=32)
    rng = np.random.default_rng(7)
    g = np.zeros(V, np.float32)
    g[rng.integers(0, V, 256)] = 1.0
    src, dst = C.DeviceArray.from_host(s), C.DeviceArray.from_host(d)
    wd, td, gd = (C.DeviceArray.from_host(x) for x in (w, t, g))
    r1, r2 = C.DeviceArray(V, np.float32), C.DeviceArray(V, np.float32)
    it = ctypes.c_uint32()
    lib = C.lib()
    assert lib.cytom_ppr_weighted(V, E, src.ptr, dst.ptr, wd.ptr, td.ptr, 0.85, 0.0, 200, r1.ptr,
                                  ctypes.byref(it)) == 0
    G = A.PPRGraph(V, s, d, w)
    G.forward_device(td, r2, tol=0.0, max_iter=200)
    assert np.array_equal(r1.to_host(), r2.to_host())
    dw1, ds1 = C.DeviceArray(E, np.float32), C.DeviceArray(V, np.float32)
    dw2, ds2 = C.DeviceArray(E, np.float32), C.DeviceArray(V, np.float32)
    assert lib.cytom_ppr_adjoint(V, E, src.ptr, dst.ptr, wd.ptr, r1.ptr, gd.ptr, 0.85, 0.0, 200,
                                 dw1.ptr, ds1.ptr, ctypes.byref(it)) == 0
    G.adjoint_device(r2, gd, dw2, ds2, tol=0.0, max_iter=200)
    assert np.array_equal(dw1.to_host(), dw2.to_host())
    assert np.array_equal(ds1.to_host(), ds2.to_host())
    w2 = (w * rng.uniform(0.5, 2.0, E)).astype(np.float32)
    G.set_weights(w2)
    assert np.array_equal(G.forward(t, tol=0.0, max_iter=200),
                          A.PPRGraph(V, s, d, w2).forward(t, tol=0.0, max_iter=200))
    G.close()
    bad = d.copy()
    bad[123] = V
    with pytest.raises(C.CytomError):
        A.PPRGraph(V, s, bad, w)
