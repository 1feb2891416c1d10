"""The CPU oracle, pinned against the reference's own code, closed forms and
textbook checkers (no GPU needed)."""
import collections
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O
from oracle import ref_build

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "rmat_kat.json")))


# ------------------------------------------------------------------ a1: hash/RMAT
def test_hash_known_answers():
    for x, h in KAT["hash64"]:
        assert O.lib().orc_hash64(x) == h
    for x, h in KAT["hash32"]:
        assert O.lib().orc_hash32(x) == h
    for x, hd in KAT["hash_double"]:
        assert O.hash_double(x) == float.fromhex(hd)


def test_rmat_known_answers():
    for case in KAT["rmat"]:
        n = len(case["src"])
        s, d = O.rmat_edges(case["scale"], case["seed"], case["first"], n)
        assert s.tolist() == case["src"], case["scale"]
        assert d.tolist() == case["dst"], case["scale"]


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference tree absent")
def test_rmat_against_reference_code():
    if not os.path.exists(ref_build.LIB):
        assert ref_build.build(quiet=True)
    R = ctypes.CDLL(ref_build.LIB)
    u32p = np.ctypeslib.ndpointer(np.uint32)
    R.ref_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                           ctypes.c_int, u32p, u32p]
    for scale in (3, 12, 22, 27):
        for rec in (0, 1):
            s = np.empty(20000, np.uint32)
            d = np.empty(20000, np.uint32)
            R.ref_rmat(scale, 2, 777, 20000, rec, s, d)
            s2, d2 = O.rmat_edges(scale, 2, 777, 20000)
            assert (s == s2).all() and (d == d2).all()


def test_rmat_iterative_equals_recursive():
    # Two structurally different restatements of the same recurrence: a wrong
    # level/stride order in either breaks equality.
    for scale in (1, 2, 7, 16, 22, 27):
        a = O.rmat_edges(scale, 2, 5, 4000)
        b = O.rmat_edges(scale, 2, 5, 4000, recursive=True)
        assert (a[0] == b[0]).all() and (a[1] == b[1]).all()


def test_rmat_quadrant_frequencies():
    # Every level picks quadrant (src bit, dst bit) = (0,0) with p=a, (0,1) b,
    # (1,0) c, (1,1) d.
    n = 200000
    s, d = O.rmat_edges(16, 2, 0, n)
    for bit in (0, 7, 15):
        sb = (s >> bit) & 1
        db = (d >> bit) & 1
        p00 = np.mean((sb == 0) & (db == 0))
        p01 = np.mean((sb == 0) & (db == 1))
        p10 = np.mean((sb == 1) & (db == 0))
        p11 = np.mean((sb == 1) & (db == 1))
        tol = 4 * np.sqrt(0.25 / n)
        assert abs(p00 - 0.57) < tol and abs(p01 - 0.19) < tol
        assert abs(p10 - 0.19) < tol and abs(p11 - 0.05) < tol


def test_weights_range():
    w = O.weights_u32(2, 0, 100000)
    assert w.min() >= 1 and w.max() <= 255
    counts = np.bincount(w, minlength=256)[1:]
    assert counts.min() > 0.7 * counts.mean()
    wf = O.weights_f32(2, 0, 100000)
    assert wf.min() > 0.0 and wf.max() <= 1.0


# ------------------------------------------------------------------ stream sampler
def test_stream_deletions_are_live_and_distinct():
    st = O.Stream(9)
    live = collections.Counter()
    for b in range(6):
        first, dels = st.next(1000, 150)
        for i in range(first, first + 1000):
            live[i] += 1
        assert len(dels) == 150
        assert len(set(dels.tolist())) == 150
        for i in dels.tolist():
            assert live[i] == 1
            live[i] -= 1
        assert st.live == sum(live.values())
    # More deletions than live indices: clamps.
    first, dels = st.next(0, 10**6)
    assert len(dels) == sum(live.values()) and st.live == 0


def test_stream_weights_are_uniform_in_1_255():
    # SSSP weights of the stream: uniform over [1, 255], consecutive weights
    # uncorrelated.
    from scipy import stats
    w = O.weights_u32(2, 0, 1 << 20)
    assert w.min() == 1 and w.max() == 255
    c = np.bincount(w, minlength=256)[1:]
    e = len(w) / 255
    chi = float(((c - e) ** 2 / e).sum())
    assert 1e-4 < stats.chi2.sf(chi, 254) < 1 - 1e-4, chi
    r = np.corrcoef(w[:-1].astype(np.float64), w[1:].astype(np.float64))[0, 1]
    assert abs(r) < 5 / np.sqrt(len(w)), r


def test_stream_deletions_are_uniform_over_live_indices():
    # The deletion draws of a batch are a uniform sample without replacement
    # of the live stream indices: over many seeds every live index is drawn
    # equally often (chi-square against the uniform expectation), also when
    # an earlier batch already removed part of the live list.  The product's
    # sampler equals this one draw for draw (test_abi).
    from scipy import stats
    n, k1, k2, seeds = 1000, 400, 100, 300
    first_batch = np.zeros(n, np.int64)
    second_batch = np.zeros(n, np.int64)
    survivors = np.zeros(n, np.int64)
    for seed in range(seeds):
        st = O.Stream(1000 + seed)
        _, d1 = st.next(n, k1)
        first_batch[d1.astype(np.int64)] += 1
        alive = np.ones(n, bool)
        alive[d1.astype(np.int64)] = False
        survivors += alive
        _, d2 = st.next(0, k2)
        assert alive[d2.astype(np.int64)].all()
        second_batch[d2.astype(np.int64)] += 1
    # Each index's count over the seeds is Binomial(seeds, k1 / n).
    p1 = k1 / n
    chi1 = float(((first_batch - seeds * p1) ** 2 / (seeds * p1 * (1 - p1))).sum())
    assert 1e-4 < stats.chi2.sf(chi1, n - 1) < 1 - 1e-4, chi1
    # Second batch: an index that survived the first batch is drawn with
    # probability k2 / (n - k1) each time it survived.
    p2 = k2 / (n - k1)
    var2 = survivors * p2 * (1 - p2)
    chi2 = float(((second_batch - survivors * p2) ** 2 / var2).sum())
    assert 1e-4 < stats.chi2.sf(chi2, n - 1) < 1 - 1e-4, chi2
    # No drift along the stream: the mean drawn index is the middle one.
    idx = np.repeat(np.arange(n), first_batch)
    assert abs(idx.mean() - (n - 1) / 2) < 5 * idx.std() / np.sqrt(len(idx))


# ------------------------------------------------------------------ store
def _multiset(s, d, w=None):
    if w is None:
        return collections.Counter(zip(s.tolist(), d.tolist()))
    return collections.Counter(zip(s.tolist(), d.tolist(), w.tolist()))


def test_store_insert_delete_matches_replay():
    rng = np.random.default_rng(3)
    V = 256
    g = O.Graph(V, vertices_per_tile=64, weighted=True)
    ref = collections.Counter()
    for b in range(5):
        s = rng.integers(0, V, 400).astype(np.uint32)
        d = rng.integers(0, V, 400).astype(np.uint32)
        w = rng.integers(1, 4, 400).astype(np.uint32)
        for e in zip(s.tolist(), d.tolist(), w.tolist()):
            ref[e] += 1
        live = list(ref.elements())
        pick = rng.choice(len(live), 120, replace=False)
        ds = np.array([live[i][0] for i in pick], np.uint32)
        dd = np.array([live[i][1] for i in pick], np.uint32)
        dw = np.array([live[i][2] for i in pick], np.uint32)
        # plus deletions that match nothing
        ds = np.concatenate([ds, np.array([1, 2], np.uint32)])
        dd = np.concatenate([dd, np.array([3, 4], np.uint32)])
        dw = np.concatenate([dw, np.array([999, 999], np.uint32)])
        found = g.update(s, d, w, ds, dd, dw)
        assert found == 120
        for i in pick:
            ref[live[i]] -= 1
        ls, ld, lw = g.live_edges()
        assert _multiset(ls, ld, lw) == +ref
        od, idg = g.degrees()
        exp_o = np.zeros(V, np.int64)
        exp_i = np.zeros(V, np.int64)
        for (a, c, _), k in ref.items():
            exp_o[a] += k
            exp_i[c] += k
        assert (od == exp_o).all() and (idg == exp_i).all()


def test_store_rejects_out_of_range():
    g = O.Graph(64, vertices_per_tile=64)
    with pytest.raises(ValueError):
        g.update(np.array([1], np.uint32), np.array([64], np.uint32))
    assert len(g.live_edges()[0]) == 0


# ------------------------------------------------------------------ executors
def _stream_run(scale, vpt, batches, n_ins, n_del, threads=1, seed=2, cap=100):
    V = 1 << scale
    st = O.Stream(seed)
    g = O.Graph(V, vertices_per_tile=vpt, weighted=True, threads=threads, max_iterations=cap)
    pids = dict(bfs=g.register(O.BFS, 0), cc=g.register(O.CC), sssp=g.register(O.SSSP, 0),
                prd=g.register(O.PRD))
    for b in range(batches):
        first, didx = st.next(n_ins, n_del if b else 0)
        s, d, w, ds, dd, dw = O.batch_edges(scale, seed, first, n_ins, didx, True)
        g.update(s, d, w, ds, dd, dw)
        stats = {k: g.execute(p) for k, p in pids.items()}
        yield b, g, pids, stats


@pytest.mark.parametrize("threads", [1, 4])
def test_executors_match_textbook_checkers(threads):
    scale = 11
    V = 1 << scale
    for b, g, pids, stats in _stream_run(scale, 512, 5, 3000, 300, threads=threads):
        ls, ld, lw = g.live_edges()
        assert (g.read(pids["bfs"]) == O.check_bfs(V, ls, ld, 0)).all()
        assert (g.read(pids["cc"]) == O.check_cc(V, ls, ld)).all()
        assert (g.read(pids["sssp"]) == O.check_sssp(V, ls, ld, lw, 0)).all()
        r = g.read(pids["prd"]).astype(np.float64)
        r64, _ = O.prd_f64(V, ls, ld)
        l1 = np.abs(r / r.sum() - r64 / r64.sum()).sum()
        assert l1 < 1e-6, l1
        for k in ("bfs", "cc", "sssp", "prd"):
            assert stats[k].converged == 1


def test_insert_only_incremental_cc_matches_checker():
    # Insert-only CC on an RMAT stream: hooks plus incremental runs, never a
    # re-execution.
    scale = 12
    V = 1 << scale
    st = O.Stream(2)
    g = O.Graph(V, vertices_per_tile=4096)
    p = g.register(O.CC)
    for b in range(8):
        first, _ = st.next(2048, 0)
        s, d = O.rmat_edges(scale, 2, first, 2048)
        g.update(s, d)
        g.execute(p)
        assert g.last_reexec(p) == 0
        ls, ld, _ = g.live_edges()
        assert (g.read(p) == O.check_cc(V, ls, ld)).all()


def test_deletion_reexecution_is_exact():
    # The counterexample to critical-neighbour re-execution: a hook lowers a
    # label through an edge that a later deletion removes.
    g = O.Graph(64, vertices_per_tile=64)
    p = g.register(O.CC)
    g.update(np.array([5], np.uint32), np.array([7], np.uint32))
    g.execute(p)
    g.update(np.array([2], np.uint32), np.array([5], np.uint32),
             dsrc=np.array([5], np.uint32), ddst=np.array([7], np.uint32))
    g.execute(p)
    lab = g.read(p)
    assert lab[7] == 7 and lab[5] == 2


def test_prd_closed_form_tiny():
    # 0 -> 1: r0 = 0.15, r1 = 0.15 + 0.85 * 0.15 = 0.2775
    g = O.Graph(64, vertices_per_tile=64)
    p = g.register(O.PRD)
    g.update(np.array([0], np.uint32), np.array([1], np.uint32))
    st = g.execute(p)
    r = g.read(p)
    assert abs(r[0] - 0.15) < 1e-7 and abs(r[1] - 0.2775) < 1e-7
    assert st.converged == 1
    # 2-cycle: r = 1 for both; isolated vertices 0.15
    g2 = O.Graph(64, vertices_per_tile=64)
    p2 = g2.register(O.PRD)
    g2.update(np.array([3, 4], np.uint32), np.array([4, 3], np.uint32))
    g2.execute(p2)
    r2 = g2.read(p2)
    assert abs(r2[3] - 1.0) < 1e-6 and abs(r2[4] - 1.0) < 1e-6
    assert abs(r2[10] - 0.15) < 1e-7


def test_prd_deleting_a_last_out_edge():
    # The DELETE hook divides by the source's out-degree with the deleted edge
    # still counted (the reference's random deletions leave the degrees
    # alone until after edgeChanged); afterwards the degrees drop.  Batch 1
    # inserts and deletes 3 -> 1, the only out-edge of 3, before any run
    # (ranks still 0): the result is the closed form of 0 -> 1, 0 -> 2.
    g = O.Graph(64, vertices_per_tile=64)
    p = g.register(O.PRD)
    u32 = lambda *x: np.array(x, np.uint32)  # noqa: E731
    assert g.update(u32(0, 0, 3), u32(1, 2, 1), dsrc=u32(3), ddst=u32(1)) == 1
    outd, ind = g.degrees()
    assert outd[3] == 0 and outd[0] == 2 and ind[1] == 1
    g.execute(p)
    r = g.read(p)
    assert np.all(np.isfinite(r))
    assert abs(r[0] - 0.15) < 1e-7 and abs(r[3] - 0.15) < 1e-7
    assert abs(r[1] - 0.21375) < 1e-6 and abs(r[2] - 0.21375) < 1e-6
    # after a run: deleting 4 -> 0, the only out-edge of 4, re-executes
    # (0.15 / 1 exceeds delta * rank[0]) and lands on the closed form again
    g.update(u32(4), u32(0))
    g.execute(p)
    g.update(u32(), u32(), dsrc=u32(4), ddst=u32(0))
    assert g.last_reexec(p) == 1
    g.execute(p)
    r = g.read(p)
    assert abs(r[1] - 0.21375) < 1e-6 and abs(r[4] - 0.15) < 1e-7


def test_prd_converges_to_power_iteration():
    scale = 10
    V = 1 << scale
    s, d = O.rmat_edges(scale, 2, 0, 8000)
    g = O.Graph(V, vertices_per_tile=256, max_iterations=1000, delta=1e-12)
    p = g.register(O.PRD)
    g.update(s, d)
    g.execute(p)
    r = g.read(p).astype(np.float64)
    pr, _ = O.pagerank_f64(V, s, d)
    assert np.abs(r - pr).max() < 1e-4
    rel = np.abs(r / r.sum() - pr / pr.sum()).sum()
    assert rel < 1e-5


def test_iteration_cap_leaves_leftovers_for_next_run():
    # A path 0->1->...->9 inserted back to front: the hooks only relax 0->1,
    # the run needs 9 iterations; with a cap of 3 it stops unconverged and
    # the next run continues from the leftover frontier.
    g = O.Graph(64, vertices_per_tile=64, max_iterations=3)
    p = g.register(O.BFS, 0)
    s = np.arange(8, -1, -1).astype(np.uint32)
    g.update(s, s + 1)
    st = g.execute(p)
    assert st.converged == 0 and st.iterations == 3
    for _ in range(3):
        g.execute(p)
    lv = g.read(p)
    assert lv[:10].tolist() == list(range(10))


# ------------------------------------------------------------ personalized
def test_ppr_f64_closed_form_cycle():
    # 3-cycle 0->1->2->0, teleport {0}: r_k = a^k (1-a) / (1 - a^3).
    a = 0.85
    s = np.array([0, 1, 2], np.uint32)
    d = np.array([1, 2, 0], np.uint32)
    r, _ = O.ppr_f64(3, s, d, [0], alpha=a)
    z = (1 - a) / (1 - a ** 3)
    assert np.allclose(r, [z, a * z, a * a * z], rtol=1e-12, atol=0)


def test_ppr_f64_uniform_teleport_is_pagerank():
    # t uniform over every vertex makes PPR the global PageRank / V.
    V = 1 << 10
    s, d = O.rmat_edges(10, 5, 0, 6000)
    pr, _ = O.pagerank_f64(V, s, d)
    ppr, _ = O.ppr_f64(V, s, d, np.arange(V, dtype=np.uint32))
    assert np.allclose(ppr * V, pr, rtol=1e-9, atol=1e-12)


def test_ppr_delta_program_tiny_and_power_iteration():
    a = 0.85
    g = O.Graph(64, vertices_per_tile=64)
    p = g.register(O.PPR, 0)
    g.update(np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32))
    g.execute(p)
    r = g.read(p)
    z = (1 - a) / (1 - a ** 3)
    assert np.allclose(r[:3], [z, a * z, a * a * z], rtol=2e-6)
    assert np.all(r[3:] == 0)
    # Random graph, several teleport sets (duplicates ignored): the FP32
    # delta program stays within its threshold of the FP64 fixpoint.
    V = 1 << 11
    s, d = O.rmat_edges(11, 9, 0, 16000)
    g = O.Graph(V, vertices_per_tile=512)
    p = g.register(O.PPR, 0)
    g.update(s, d)
    for tel in ([7], [1, 2, 3, 3, 900], list(range(0, V, 17))):
        g.set_sources(p, tel)
        st = g.execute(p)
        assert st.converged == 1
        r = g.read(p).astype(np.float64)
        ref, _ = O.ppr_f64(V, s, d, tel)
        err = np.abs(r - ref).sum() / ref.sum()
        assert err < 2e-5, (tel[:3], err)


# ------------------------------------------------- differentiable PPR (FP64)
def _weighted_graph(V, E, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, V, E).astype(np.uint32)
    d = rng.integers(0, V, E).astype(np.uint32)
    w = rng.uniform(1e-3, 1.0, E)
    t = np.zeros(V)
    t[rng.choice(V, 16, replace=False)] = 1.0 / 16
    return s, d, w, t


def test_ppr_grad_oracle_dense_equals_iterative():
    from oracle import ppr_grad as G
    V = 256
    s, d, w, t = _weighted_graph(V, 2000, 1)
    r_it = G.forward(V, s, d, w, t)
    r_de = G.forward(V, s, d, w, t, dense=True)
    assert np.allclose(r_it, r_de, rtol=1e-11, atol=1e-14)
    _, g = G.loss_sq(r_de, np.arange(0, V, 3))
    dw_it, ds_it, _ = G.adjoint(V, s, d, w, r_de, g)
    dw_de, ds_de, _ = G.adjoint(V, s, d, w, r_de, g, dense=True)
    assert np.allclose(dw_it, dw_de, rtol=1e-9, atol=1e-14)
    assert np.allclose(ds_it, ds_de, rtol=1e-9, atol=1e-14)


def test_ppr_grad_oracle_matches_finite_differences():
    # Central differences, h = 1e-3, on sampled edges and teleport entries.
    from oracle import ppr_grad as G
    V = 512
    s, d, w, t = _weighted_graph(V, 4000, 2)
    targets = np.random.default_rng(3).integers(0, V, 64)

    def L(wv, tv):
        return G.loss_sq(G.forward(V, s, d, wv, tv, dense=True), targets)[0]

    r = G.forward(V, s, d, w, t, dense=True)
    _, g = G.loss_sq(r, targets)
    dw, ds, _ = G.adjoint(V, s, d, w, r, g, dense=True)
    h = 1e-3
    for e in np.random.default_rng(4).choice(len(w), 24, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (L(wp, t) - L(wm, t)) / (2 * h)
        assert abs(fd - dw[e]) <= 1e-6 * max(1.0, abs(dw[e])) + 1e-5 * abs(dw[e]), (e, fd, dw[e])
    for v in (0, 5, 77):
        tp, tm = t.copy(), t.copy()
        tp[v] += h
        tm[v] -= h
        fd = (L(w, tp) - L(w, tm)) / (2 * h)
        assert abs(fd - ds[v]) <= 1e-9 + 1e-6 * abs(ds[v])


def test_ppr_grad_oracle_uniform_teleport_is_pagerank():
    from oracle import ppr_grad as G
    V = 1 << 9
    s, d = O.rmat_edges(9, 3, 0, 3000)
    r = G.forward(V, s, d, np.ones(len(s)), np.full(V, 1.0 / V))
    pr, _ = O.pagerank_f64(V, s, d)
    assert np.allclose(r * V, pr, rtol=1e-9)
