"""Parity at the sizes of the measured workloads, through the C-ABI.

* RMAT-22 stream (64 batches of 943,719 inserts + 104,857 deletions, BFS /
  CC / SSSP / PageRank-delta after every batch): every 8 batches the store
  equals a host replay of the stream (multiset and degrees), BFS / CC / SSSP
  equal the textbook checkers, and at three of those batches PageRank-delta
  is within 1e-6 normalised L1 of the FP32 oracle (error against the FP64
  restatement printed and bounded).
* Batched queries on the final RMAT-22 graph: 256 BFS + 256 SSSP + 64
  personalized PageRank in one call; 16 of each against the textbook
  checkers and against single-query runs.
* 128 small evolving graphs as one union context: every tenant's CC labels
  against the textbook checker after batches 4 and 8.
* Differentiable PPR on RMAT-20: adjoint against the FP64 adjoint and central
  differences (h = 1e-3) on 512 edges.
* RMAT-27 stream (128 batches of 2^24 updates, 5 % deletions), CC after every
  batch: labels against the textbook checker after batches 64 and 128, store
  degrees against the exported edges, final PageRank-delta by its FP64
  fixed-point residual.

Long tests print progress lines (capture disabled for them) so a slow check
is never silent for long.
"""
import time

import numpy as np
import pytest

import cytom_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

PRD_L1_TOL = 1e-6


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert C.device_count() > 0, "no HIP device visible: the GPU tests need an MI355X"
    O.checker_threads(O.host_threads())


@pytest.fixture
def say(capsys):
    t0 = time.perf_counter()

    def f(msg):
        with capsys.disabled():
            print("  [%6.1fs] %s" % (time.perf_counter() - t0, msg), flush=True)

    return f


def l1n(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def sorted_keys(s, d, w):
    """(src, dst, weight) triples of ids < 2^22 and weights < 2^8 as sorted
    uint64 keys: equal arrays <=> equal multisets."""
    k = (s.astype(np.uint64) << np.uint64(30)) | (d.astype(np.uint64) << np.uint64(8))
    k |= w.astype(np.uint64)
    k.sort()
    return k


def chunked_bincount(a, V, chunk=1 << 27):
    out = np.zeros(V, np.int64)
    for i in range(0, len(a), chunk):
        out += np.bincount(a[i:i + chunk], minlength=V)
    return out


# --------------------------------------------------------------- RMAT-22 stream
@pytest.mark.timeout(900)
def test_c2_rmat22_stream_checked_every_8_batches(say):
    scale, V, n_ins, n_del, seed = 22, 1 << 22, 943719, 104857, 2
    threads = O.host_threads()
    st = C.Stream(seed)
    g = C.Graph(V, weighted=True, max_iterations=100, delta=1e-7, edge_capacity=3 * (1 << 26))
    names = ("bfs", "cc", "sssp", "pagerank-delta")
    ids = {n: g.register(n, 0) for n in names}
    total = 64 * n_ins
    # Host replay: every inserted stream index with its oracle-generated edge,
    # and whether it is still live.
    hs = np.empty(total, np.uint32)
    hd = np.empty(total, np.uint32)
    hw = np.empty(total, np.uint32)
    alive = np.zeros(total, bool)
    prd_checks = (24, 48, 64)
    for b in range(64):
        first, didx = st.next(n_ins, n_del)
        hs[first:first + n_ins], hd[first:first + n_ins] = O.rmat_edges(scale, seed, first, n_ins)
        hw[first:first + n_ins] = O.weights_u32(seed, first, n_ins)
        alive[first:first + n_ins] = True
        alive[didx.astype(np.int64)] = False
        s, d, w = C.rmat_generate(scale, seed, first, n_ins, weights=True)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(scale, seed, di, len(didx), weights=True)
        bs = g.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
        assert bs.deleted == len(didx) and bs.missing == 0, b  # drawn from live edges
        runs = {n: g.execute(pid) for n, pid in ids.items()}
        for n in ("bfs", "cc", "sssp"):
            assert runs[n].converged == 1, (b, n)
        if (b + 1) % 8:
            continue
        say("rmat22 stream: checks after batch %d (%d live edges)" % (b + 1, bs.live_edges))
        ls, ld, lw = g.export_edges()
        assert len(ls) == bs.live_edges == int(alive.sum()), b
        rs, rd, rw = hs[alive], hd[alive], hw[alive]
        assert np.array_equal(sorted_keys(ls, ld, lw), sorted_keys(rs, rd, rw)), b
        go, gi = g.degrees()
        assert np.array_equal(go, np.bincount(rs, minlength=V)), b
        assert np.array_equal(gi, np.bincount(rd, minlength=V)), b
        del rs, rd, rw
        assert np.array_equal(g.read(ids["bfs"]), O.check_bfs(V, ls, ld, 0)), b
        assert np.array_equal(g.read(ids["cc"]), O.check_cc(V, ls, ld, threads)), b
        assert np.array_equal(g.read(ids["sssp"]), O.check_sssp(V, ls, ld, lw, 0)), b
        if b + 1 in prd_checks:
            prd = ids["pagerank-delta"]
            if not runs["pagerank-delta"].reexecuted:  # compare runs from init
                g.init(prd)
                g.execute(prd)
            ranks = g.read(prd)
            og = O.Graph(V, vertices_per_tile=16384, max_iterations=100, delta=1e-7,
                         threads=threads)
            og.update(ls, ld)
            op = og.register(O.PRD)
            og.execute(op)
            r64, _ = O.prd_f64(V, ls, ld, threads=threads)
            e32, e64, eo = l1n(ranks, og.read(op)), l1n(ranks, r64), l1n(og.read(op), r64)
            say("rmat22 stream: PageRank-delta after batch %d: L1 vs FP32 oracle %.3g, vs FP64 "
                "%.3g (oracle FP32 vs FP64 %.3g)" % (b + 1, e32, e64, eo))
            # The GPU's neighbour sums are exact (int64 fixed point); the FP32
            # oracle's float sums round in thread order, which at 40 M edges
            # alone puts it ~1e-6 from FP64.  The bound: the GPU within 1e-6
            # of the FP64 restatement, and within 1e-6 plus the oracle's own
            # FP64 distance of the FP32 oracle (triangle inequality).
            assert e64 <= PRD_L1_TOL, (b, e64)
            assert e32 <= PRD_L1_TOL + eo, (b, e32, eo)
            del og
    assert bs.live_edges == 64 * (n_ins - n_del)


# ----------------------------------------------------------- batched queries
@pytest.mark.timeout(900)
def test_c3_batched_queries_on_the_final_rmat22_graph(say):
    scale, V, n_ins, n_del, seed = 22, 1 << 22, 943719, 104857, 2
    threads = O.host_threads()
    st = C.Stream(seed)
    g = C.Graph(V, weighted=True, max_iterations=100, delta=1e-7, edge_capacity=3 * (1 << 26))
    for b in range(64):
        first, didx = st.next(n_ins, n_del)
        s, d, w = C.rmat_generate(scale, seed, first, n_ins, weights=True)
        di = C.DeviceArray.from_host(didx)
        ds, dd, dw = C.rmat_gather(scale, seed, di, len(didx), weights=True)
        g.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
    say("c3: stream of 64 batches ingested (%d live edges)" % g.live_edge_count())
    deg_o, deg_i = g.degrees()
    live = np.nonzero(deg_o.astype(np.int64) + deg_i)[0].astype(np.uint32)
    rng = np.random.default_rng(7)
    bfs_src = rng.choice(live, 256)
    sssp_src = rng.choice(live, 256)
    ppr_sets = [rng.choice(live, int(rng.integers(1, 1025))) for _ in range(64)]
    queries = ([("bfs", int(x)) for x in bfs_src] + [("sssp", int(x)) for x in sssp_src] +
               [("ppr", t) for t in ppr_sets])
    order = rng.permutation(len(queries))
    queries = [queries[i] for i in order]
    t0 = time.perf_counter()
    rs = g.batched_execute(queries)
    say("c3: %d queries batched in %.1f ms (%d traversed edges)"
        % (len(queries), (time.perf_counter() - t0) * 1e3, rs.traversed_edges))
    picks = {k: [i for i, q in enumerate(queries) if q[0] == k][:16] for k in ("bfs", "sssp", "ppr")}
    ls, ld, lw = g.export_edges()
    ref_bfs = O.check_multi(V, ls, ld, [queries[i][1] for i in picks["bfs"]], threads=threads)
    ref_sssp = O.check_multi(V, ls, ld, [queries[i][1] for i in picks["sssp"]], w=lw,
                             threads=threads)
    say("c3: textbook BFS and Dijkstra from 16 sources each done")
    for k, i in enumerate(picks["bfs"]):
        assert np.array_equal(g.read_query(i), ref_bfs[k]), ("bfs", i)
    for k, i in enumerate(picks["sssp"]):
        assert np.array_equal(g.read_query(i), ref_sssp[k]), ("sssp", i)
    pid = {k: g.register(k, int(live[0])) for k in ("bfs", "sssp", "ppr")}
    for kind, idx in picks.items():
        for i in idx:
            g.set_sources(pid[kind], queries[i][1])
            g.execute(pid[kind])
            single = g.read(pid[kind])
            got = g.read_query(i)
            if kind == "ppr":
                assert l1n(got, single) <= PRD_L1_TOL, (kind, i)
            else:
                assert np.array_equal(got, single), (kind, i)
    say("c3: 48 queries equal their single-query runs")


@pytest.mark.timeout(900)
def test_c3_tenant_union_cc_against_checker(say):
    # 128 graphs, scale U{10..20}, edge factor U{8..32}, each its own RMAT
    # stream cut into 8 insert batches, held side by side in one context.
    rng = np.random.default_rng(202)
    specs = [(int(rng.integers(10, 21)), int(rng.integers(8, 33))) for _ in range(128)]
    bases = np.cumsum([0] + [1 << sc for sc, _ in specs])
    V_union = int(bases[-1])
    nb = 8
    G = C.Graph(V_union, edge_capacity=int(sum(ef << sc for sc, ef in specs) * 1.3) + 1024)
    p = G.register("cc")
    hist = [[] for _ in specs]  # host copies of each tenant's edges so far
    for b in range(nb):
        us, ud = [], []
        for gi, (sc, ef) in enumerate(specs):
            e_b = (ef << sc) // nb
            s, d, _ = C.rmat_generate(sc, 3000 + gi, b * e_b, e_b)
            hs, hd = s.to_host(), d.to_host()
            hist[gi].append((hs, hd))
            us.append(hs + np.uint32(bases[gi]))
            ud.append(hd + np.uint32(bases[gi]))
        us, ud = np.concatenate(us), np.concatenate(ud)
        G.update_device(C.DeviceArray.from_host(us), C.DeviceArray.from_host(ud), None, len(us))
        r = G.execute(p)
        assert r.converged == 1, b
        if b + 1 not in (4, 8):
            continue
        lab = G.read(p)
        for gi, (sc, _) in enumerate(specs):
            s = np.concatenate([x[0] for x in hist[gi]])
            d = np.concatenate([x[1] for x in hist[gi]])
            part = lab[bases[gi]:bases[gi] + (1 << sc)].astype(np.int64) - int(bases[gi])
            assert np.array_equal(part, O.check_cc(1 << sc, s, d).astype(np.int64)), (b, gi)
        say("tenants: CC of all 128 tenants equal the checker after batch %d (%d edges)"
            % (b + 1, G.live_edge_count()))


# ------------------------------------------------ differentiable PPR (C4 size)
@pytest.mark.timeout(900)
def test_c4_ppr_grad_rmat20_fp64_and_finite_differences(say):
    from cytom_amd import autograd as A
    from oracle import ppr_grad as PG
    scale, V, E = 20, 1 << 20, 1 << 24
    s, d = O.rmat_edges(scale, 20, 0, E)
    rng = np.random.default_rng(20)
    w = (1.0 - rng.random(E)).astype(np.float32)  # U(0, 1]
    t = np.zeros(V, np.float32)
    t[rng.choice(V, 256, replace=False)] = 1.0 / 256
    targets = rng.integers(0, V, 1024)
    G = A.PPRGraph(V, s, d, w)
    r = G.forward(t)
    g = np.zeros(V, np.float32)
    np.add.at(g, targets, 2.0 * r[targets])
    dw, ds = G.adjoint(r, g)
    r_ref = PG.forward(V, s, d, w.astype(np.float64), t.astype(np.float64), tol=1e-13)
    _, g_ref = PG.loss_sq(r_ref, targets)
    dw_ref, ds_ref, _ = PG.adjoint(V, s, d, w.astype(np.float64), r_ref, g_ref, tol=1e-13)

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    er, edw, eds = rel(r, r_ref), rel(dw, dw_ref), rel(ds, ds_ref)
    say("c4: relative error vs FP64: r %.2e, dL/dw %.2e, dL/ds %.2e" % (er, edw, eds))
    assert er < 1e-5 and edw < 1e-4 and eds < 1e-4
    # Central differences on 512 edges drawn with probability ~ |dL/dw|
    # (most gradients sit below the FP32 resolution of the loss).  Fixed
    # sweep count so both sides of a difference run the same iteration.
    pw = np.abs(dw.astype(np.float64))
    picks = rng.choice(E, 512, replace=False, p=pw / pw.sum())
    sweeps = 160

    def loss(wv):
        G.set_weights(wv)
        rv = G.forward(t, tol=0.0, max_iter=sweeps).astype(np.float64)
        return float(np.sum(rv[targets] ** 2))

    h = 1e-3
    fd = np.empty(len(picks))
    for k, e in enumerate(picks):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd[k] = (loss(wp) - loss(wm)) / (2 * h)
        if k % 128 == 127:
            say("c4: %d central differences" % (k + 1))
    efd = rel(dw[picks], fd)
    efd64 = rel(dw_ref[picks], fd)
    say("c4: dL/dw on 512 edges: GPU adjoint vs central differences %.2e (FP64 adjoint vs the "
        "same differences %.2e)" % (efd, efd64))
    assert efd < 2e-2
    G.close()


# --------------------------------------------------------------- RMAT-27 stream
@pytest.mark.timeout(1200)
def test_c5_rmat27_stream(say):
    scale, V, per, seed = 27, 1 << 27, 1 << 24, 2
    n_del = 838860
    n_ins = per - n_del
    threads = O.host_threads()
    g = C.Graph(V, max_iterations=100, delta=1e-7, edge_capacity=3 * (1 << 31))
    g.set_tuning(chunk_slots=64)
    cc = g.register("cc")
    prd = g.register("pagerank-delta")
    st = C.Stream(seed)
    t_cc = 0.0
    for b in range(128):
        first, didx = st.next(n_ins, n_del)
        s, d, _ = C.rmat_generate(scale, seed, first, n_ins)
        di = C.DeviceArray.from_host(didx)
        ds, dd, _ = C.rmat_gather(scale, seed, di, len(didx))
        bs = g.update_device(s, d, None, n_ins, ds, dd, None, len(didx))
        assert bs.deleted == len(didx) and bs.missing == 0, b
        del s, d, di, ds, dd
        r = g.execute(cc)
        t_cc += r.t_compute_ms
        assert r.converged == 1, b
        if (b + 1) % 16 == 0:
            say("rmat27: batch %d, %d live edges, CC %.1f ms/batch"
                % (b + 1, bs.live_edges, t_cc / (b + 1)))
        if b + 1 not in (64, 128):
            continue
        ls, ld, _ = g.export_edges(weights=False)
        assert len(ls) == bs.live_edges == (b + 1) * (n_ins - n_del), b
        go, gi = g.degrees()
        assert int(go.sum()) == int(gi.sum()) == len(ls)
        assert np.array_equal(go, chunked_bincount(ls, V)), b
        assert np.array_equal(gi, chunked_bincount(ld, V)), b
        say("rmat27: store degrees equal the %d exported edges after batch %d" % (len(ls), b + 1))
        assert np.array_equal(g.read(cc), O.check_cc(V, ls, ld, threads)), b
        say("rmat27: CC labels equal the checker after batch %d" % (b + 1))
        if b + 1 == 128:
            rp = g.execute(prd)  # from init on the final graph (see below)
            ranks = g.read(prd)
            assert np.isfinite(ranks).all()
            res = O.prd_residual(V, ls, ld, ranks, threads=threads)
            say("rmat27: PageRank-delta %d iterations, FP64 fixed-point residual %.3g "
                "(normalised L1)" % (rp.iterations, res))
            # A run that drops deltas below 1e-7 * rank leaves at most
            # 0.85e-7 of normalised residual; FP32 rank sums add a few 1e-7.
            assert res <= 1e-6
        del ls, ld
