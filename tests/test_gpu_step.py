"""One-call steps (cytom_step): the update batch and every program's first
iteration group are enqueued before the host reads anything.

A step's values must equal those of update_device followed by execute of
each program (bit-exact for all four programs: the PageRank-delta sums are
exact fixed point, so its trajectory is identical too) and the oracle's; a
refused batch must change nothing and run nothing; a batch that outgrows the
pool takes the repack path and still agrees.
"""
import numpy as np
import pytest

import cytom_amd as C
import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("bfs", "cc", "sssp", "pagerank-delta")
ORACLE_KINDS = (O.BFS, O.CC, O.SSSP, O.PRD)


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert C.device_count() > 0, "no HIP device visible: the GPU tests need an MI355X"


def l1n(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a / a.sum() - b / b.sum()).sum())


def make(V, vpt, cap=0):
    g = C.Graph(V, vertices_per_tile=vpt, weighted=True, edge_capacity=cap)
    return g, [g.register(n, 0) for n in NAMES]


def device_batch(scale, seed, first, n_ins, didx):
    s, d, w = C.rmat_generate(scale, seed, first, n_ins, weights=True)
    di = C.DeviceArray.from_host(didx)
    ds, dd, dw = C.rmat_gather(scale, seed, di, len(didx), weights=True)
    return s, d, w, ds, dd, dw


@pytest.mark.parametrize("cap", [0, 4096])
def test_step_equals_separate_calls_and_oracle(cap):
    # cap 4096: a pool far below the graph, so steps overflow it and take
    # the repack path.
    scale, vpt, seed, n_ins, n_del = 14, 1024, 3, 20000, 2000
    V = 1 << scale
    st = O.Stream(seed)
    og = O.Graph(V, vertices_per_tile=vpt, weighted=True)
    oids = [og.register(k, 0) for k in ORACLE_KINDS]
    ga, ia = make(V, vpt, cap)
    gb, ib = make(V, vpt, cap)
    syncs = []
    for b in range(8):
        first, didx = st.next(n_ins, n_del if b else 0)
        s, d, w, ds, dd, dw = device_batch(scale, seed, first, n_ins, didx)
        ba = ga.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
        ra = [ga.execute(p) for p in ia]
        bb, rb = gb.step(ib, s, d, w, n_ins, ds, dd, dw, len(didx))
        assert (bb.inserted, bb.deleted, bb.missing, bb.live_edges) == \
            (ba.inserted, ba.deleted, ba.missing, ba.live_edges), b
        syncs.append(bb.host_syncs)
        hs, hd, hw, hds, hdd, hdw = O.batch_edges(scale, seed, first, n_ins, didx, True)
        og.update(hs, hd, hw, hds, hdd, hdw)
        for k, name in enumerate(NAMES):
            va, vb = ga.read(ia[k]), gb.read(ib[k])
            assert np.array_equal(va, vb), (b, name)
            assert rb[k].reexecuted == ra[k].reexecuted, (b, name)
            assert rb[k].converged == ra[k].converged == 1, (b, name)
            og.execute(oids[k])
            vo = og.read(oids[k])
            if name == "pagerank-delta":
                assert (rb[k].iterations, rb[k].traversed_edges) == \
                    (ra[k].iterations, ra[k].traversed_edges), b
                assert l1n(vb, vo) <= 1e-6, b
            else:
                assert np.array_equal(vb, vo), (b, name)
    if cap == 0:
        # From the third batch on, every program's first group covers its run
        # or one more group does: at most two synchronisations per step.
        assert max(syncs[2:]) <= 2, syncs


def test_step_subset_and_argument_errors():
    scale, vpt, seed = 12, 256, 5
    V = 1 << scale
    g, ids = make(V, vpt)
    h, hid = make(V, vpt)
    st = O.Stream(seed)
    first, didx = st.next(5000, 0)
    s, d, w, ds, dd, dw = device_batch(scale, seed, first, 5000, didx)
    # a subset of the programs, in another order
    bs, rs = g.step([ids[3], ids[1]], s, d, w, 5000)
    h.update_device(s, d, w, 5000)
    for k in (3, 1):
        h.execute(hid[k])
        assert np.array_equal(g.read(ids[k]), h.read(hid[k]))
    assert len(rs) == 2 and rs[0].iterations > 0
    with pytest.raises(C.CytomError):
        g.step([ids[0], ids[0]], s, d, w, 5000)
    with pytest.raises(C.CytomError):
        g.step([ids[0], 99], s, d, w, 5000)
    # neither refused call changed the store
    assert g.live_edge_count() == h.live_edge_count()


def test_step_refused_batch_changes_nothing():
    scale, vpt, seed = 12, 256, 7
    V = 1 << scale
    g, ids = make(V, vpt)
    h, hid = make(V, vpt)
    st = O.Stream(seed)
    batches = []
    for b in range(3):
        first, didx = st.next(6000, 600 if b else 0)
        batches.append(device_batch(scale, seed, first, 6000, didx) + (len(didx),))
    s, d, w, ds, dd, dw, nd = batches[0]
    g.step(ids, s, d, w, 6000, ds, dd, dw, nd)
    h.step(hid, s, d, w, 6000, ds, dd, dw, nd)
    before = [g.read(p) for p in ids]
    live = g.live_edge_count()
    # an insertion with an out-of-range vertex refuses the whole batch
    bad = s.to_host()
    bad[17] = V
    bad_src = C.DeviceArray.from_host(bad)
    s1, d1, w1, ds1, dd1, dw1, nd1 = batches[1]
    with pytest.raises(C.CytomError) as e:
        g.step(ids, bad_src, d1, w1, 6000, ds1, dd1, dw1, nd1)
    assert e.value.code == -22
    assert g.live_edge_count() == live
    for p, v in zip(ids, before):
        assert np.array_equal(g.read(p), v)
    # the next valid steps agree with a graph that never saw the refused one
    for s, d, w, ds, dd, dw, nd in batches[1:]:
        g.step(ids, s, d, w, 6000, ds, dd, dw, nd)
        h.step(hid, s, d, w, 6000, ds, dd, dw, nd)
        for p, q in zip(ids, hid):
            assert np.array_equal(g.read(p), h.read(q))


@pytest.mark.parametrize("cap", [0, 4096])
def test_step_graphs_equal_direct_steps(cap):
    # Steps through recorded graphs against direct steps on the same stream.
    # The updates are staged in fixed device buffers (a step graph is keyed
    # by its argument pointers), batch sizes change once mid-stream (a new
    # graph) and, with cap 4096, the pool overflows and is repacked (new
    # graphs after each repack).
    scale, vpt, seed = 14, 1024, 4
    V = 1 << scale
    st = O.Stream(seed)
    ga, ia = make(V, vpt, cap)
    gb, ib = make(V, vpt, cap)
    gb.set_step_graphs(True)
    n_max = 20000
    bufs = [C.DeviceArray(n_max, np.uint32) for _ in range(6)]
    sizes = [(20000, 2000)] * 6 + [(12000, 1500)] * 6
    for b, (n_ins, n_del) in enumerate(sizes):
        first, didx = st.next(n_ins, n_del if b else 0)
        arrs = device_batch(scale, seed, first, n_ins, didx)
        for buf, a in zip(bufs, arrs):
            buf.copy_from(a)
        s, d, w, ds, dd, dw = bufs
        ba, ra = ga.step(ia, s, d, w, n_ins, ds, dd, dw, len(didx))
        bb, rb = gb.step(ib, s, d, w, n_ins, ds, dd, dw, len(didx))
        assert (bb.inserted, bb.deleted, bb.missing, bb.live_edges) == \
            (ba.inserted, ba.deleted, ba.missing, ba.live_edges), b
        for k, name in enumerate(NAMES):
            assert np.array_equal(ga.read(ia[k]), gb.read(ib[k])), (b, name)
            assert (rb[k].converged, rb[k].reexecuted) == (ra[k].converged, ra[k].reexecuted), \
                (b, name)
            if name == "pagerank-delta":
                # Its trajectory is exact (fixed-point sums); the min kinds'
                # insert hooks race on purpose (any order reaches the same
                # fixpoint), so only their values are compared.
                assert (rb[k].iterations, rb[k].traversed_edges) == \
                    (ra[k].iterations, ra[k].traversed_edges), b
    captures, replays, kept = gb.step_graph_stats()
    assert captures >= 2 and kept <= 32, (captures, replays, kept)
    if cap == 0:
        # Equal batches with the same buffers replay after the first few
        # batches (scratch buffers stop growing, parities repeat).
        assert replays >= 4, (captures, replays)
    assert ga.step_graph_stats() == (0, 0, 0)
    gb.set_step_graphs(False)
    assert gb.step_graph_stats()[2] == 0


def test_concurrent_runs_equal_serial_runs():
    # The same stream with the programs run concurrently (one stream each)
    # and one after the other: identical values, and for PageRank-delta an
    # identical trajectory.
    scale, vpt, seed, n_ins, n_del = 14, 512, 6, 16000, 1600
    V = 1 << scale
    st = O.Stream(seed)
    ga, ia = make(V, vpt)
    gb, ib = make(V, vpt)
    ga.set_concurrent_runs(False)
    gb.set_concurrent_runs(True)
    for b in range(6):
        first, didx = st.next(n_ins, n_del if b else 0)
        s, d, w, ds, dd, dw = device_batch(scale, seed, first, n_ins, didx)
        ba, ra = ga.step(ia, s, d, w, n_ins, ds, dd, dw, len(didx))
        bb, rb = gb.step(ib, s, d, w, n_ins, ds, dd, dw, len(didx))
        assert (bb.deleted, bb.live_edges) == (ba.deleted, ba.live_edges), b
        for k, name in enumerate(NAMES):
            assert np.array_equal(ga.read(ia[k]), gb.read(ib[k])), (b, name)
            assert rb[k].converged == ra[k].converged == 1, (b, name)
        k = NAMES.index("pagerank-delta")
        assert (rb[k].iterations, rb[k].traversed_edges) == \
            (ra[k].iterations, ra[k].traversed_edges), b


def test_sampled_profiling_times_every_nth_launch():
    # Profiling every launch and every 4th launch of the same runs: the
    # sampled profile holds every 4th iteration of each run, and a negative
    # period is refused.
    scale, vpt, seed = 13, 512, 10
    V = 1 << scale
    st = O.Stream(seed)
    g, ids = make(V, vpt)
    first, didx = st.next(12000, 0)
    s, d, w, ds, dd, dw = device_batch(scale, seed, first, 12000, didx)
    g.update_device(s, d, w, 12000, ds, dd, dw, len(didx))
    prd = ids[NAMES.index("pagerank-delta")]
    counts = {}
    for every in (1, 4):
        g.init(prd)
        g.set_profiling(True, every=every)
        r = g.execute(prd)
        ms, launches, alg = g.kernel_profile(C.PRD)
        counts[every] = (r.iterations, launches, ms, alg)
        g.set_profiling(False)
    it1, n1, ms1, alg1 = counts[1]
    it4, n4, ms4, alg4 = counts[4]
    assert it1 == it4 > 8
    assert n1 == it1 and it4 // 4 <= n4 <= (it4 + 3) // 4
    assert 0 < ms4 < ms1 and 0 < alg4 < alg1
    with pytest.raises(C.CytomError):
        g.set_profiling(True, every=-1)


def test_execute_many_equals_separate_runs():
    # Several programs run together (concurrently) on the same graph equal
    # their separate runs, including after re-initialisation and with a
    # program listed alone; argument errors are refused.
    scale, vpt, seed, n_ins, n_del = 13, 512, 8, 12000, 1200
    V = 1 << scale
    st = O.Stream(seed)
    ga, ia = make(V, vpt)
    gb, ib = make(V, vpt)
    for b in range(4):
        first, didx = st.next(n_ins, n_del if b else 0)
        s, d, w, ds, dd, dw = device_batch(scale, seed, first, n_ins, didx)
        ga.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
        gb.update_device(s, d, w, n_ins, ds, dd, dw, len(didx))
        if b == 2:
            for g, ids in ((ga, ia), (gb, ib)):
                for p in ids:
                    g.init(p)
        ra = [ga.execute(p) for p in ia]
        rb = gb.execute_many(ib)
        for k, name in enumerate(NAMES):
            assert np.array_equal(ga.read(ia[k]), gb.read(ib[k])), (b, name)
            assert rb[k].converged == ra[k].converged == 1, (b, name)
    [r] = gb.execute_many([ib[3]])
    assert r.converged == 1
    with pytest.raises(C.CytomError):
        gb.execute_many([ib[0], ib[0]])
    with pytest.raises(C.CytomError):
        gb.execute_many([len(NAMES) + 5])
