"""Reverse mode of a registered personalized-PageRank program on an evolving
graph (cytom_ppr_program_adjoint) against the FP64 oracle on the exported
live edge list with every weight 1, and against the explicit-edge-list
adjoint (cytom_ppr_graph_adjoint) on the same list."""
import numpy as np
import pytest

import cytom_amd as C
import oracle as O
from oracle import ppr_grad as PG

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert C.device_count() > 0, "no HIP device visible: the GPU tests need an MI355X"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("scale,vpt", [(12, 256), (16, 4096)])
def test_program_adjoint_on_an_evolving_graph(scale, vpt):
    from cytom_amd import autograd as A
    seed = 9
    V = 1 << scale
    rng = np.random.default_rng(scale)
    g = C.Graph(V, vertices_per_tile=vpt, weighted=False, max_iterations=400, delta=1e-10)
    pid = g.register("ppr", 0)
    tele = rng.choice(V, 7, replace=False).astype(np.uint32)
    g.set_sources(pid, tele)
    bfs = g.register("bfs", 0)
    st = O.Stream(seed)
    n = 8 * V
    for b in range(3):
        first, didx = st.next(n, n // 10 if b else 0)
        s, d, _, ds, dd, _ = O.batch_edges(scale, seed, first, n, didx, False)
        g.update(s, d, None, ds, dd, None)
        g.execute(pid)
    r = g.read(pid)
    es, ed, _ = g.export_edges(weights=False)
    ones = np.ones(len(es))
    t = np.zeros(V)
    t[tele] = 1.0 / len(tele)
    # The program's FP32 ranks sum some 150 deltas each (delta 1e-10): their
    # distance to the FP64 fixed point is FP32 accumulation, not the adjoint.
    r64 = PG.forward(V, es, ed, ones, t, tol=1e-15)
    assert rel(r, r64) < 1e-4
    targets = rng.integers(0, V, 64)
    gr = np.zeros(V, np.float32)
    np.add.at(gr, targets, 2.0 * r[targets])
    dt, ps, pd, dw, sweeps = g.ppr_adjoint(pid, gr)
    # tol 0: sweeps until an FP32 fixed point (no change) or the 400 cap
    assert 20 < sweeps <= 400
    # a tolerance stops the sweeps earlier
    assert g.ppr_adjoint(pid, gr, tol=1e-6)[4] < sweeps
    assert np.array_equal(ps, es) and np.array_equal(pd, ed)
    # FP64 adjoint at the program's own ranks and loss gradient
    dw64, dt64, _ = PG.adjoint(V, es, ed, ones, r.astype(np.float64), gr.astype(np.float64),
                               tol=1e-15)
    assert rel(dw, dw64) < 1e-4, rel(dw, dw64)
    assert rel(dt, dt64) < 1e-4, rel(dt, dt64)
    # the explicit-edge-list operator with the same edges and unit weights
    G = A.PPRGraph(V, es, ed, np.ones(len(es), np.float32))
    dw2, dt2 = G.adjoint(r, gr)
    assert rel(dw, dw2) < 1e-5 and rel(dt, dt2) < 1e-5
    G.close()
    # another kind of program, and an edge output below the live edges
    with pytest.raises(C.CytomError):
        g.ppr_adjoint(bfs, gr)
    E = len(es)
    out = C.DeviceArray(E - 1, np.float32)
    gd = C.DeviceArray.from_host(gr)
    rc = C.lib().cytom_ppr_program_adjoint(g._h, pid, gd.ptr, 1e-9, 0, None, None, None,
                                           out.ptr, E - 1, None)
    assert rc == -34
