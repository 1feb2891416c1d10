"""The allocation-free weighted PPR forward (cytom_ppr_weighted_noalloc)
against the allocating one-shot entry point (cytom_ppr_weighted), through the
C-ABI: bit-identical ranks and sweep counts on two calls in a row, the second
on a smaller graph (the scratch keeps the first call's contents), and a graph
beyond the scratch bound refused before any work."""
import ctypes

import numpy as np
import pytest

import cytom_amd as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_device():
    assert C.device_count() > 0, "no HIP device visible: the GPU tests need an MI355X"


def _rmat_graph(scale, E, seed, n_tel):
    s, d, _ = C.rmat_generate(scale, seed, 0, E)  # skewed: in-degree hubs take long rows
    V = 1 << scale
    rng = np.random.default_rng(seed)
    w = (1.0 - rng.random(E)).astype(np.float32)
    t = np.zeros(V, np.float32)
    t[rng.choice(V, n_tel, replace=False)] = 1.0 / n_tel
    return V, s.to_host(), d.to_host(), w, t


def _forward(fn, V, s, d, w, t, tol, max_iter):
    dev = [C.DeviceArray.from_host(x) for x in (s, d, w, t)]
    r = C.DeviceArray(V, np.float32)
    it = ctypes.c_uint32(0)
    rc = fn(V, len(s), dev[0].ptr, dev[1].ptr, dev[2].ptr, dev[3].ptr, 0.85, tol, max_iter, r.ptr,
            ctypes.byref(it))
    return rc, (r.to_host() if rc == 0 else None), it.value


def test_noalloc_forward_equals_the_allocating_forward_twice():
    lib = C.lib()
    cases = [(_rmat_graph(18, 1 << 20, 5, 256), 0.0, 120),   # near the scratch bound
             (_rmat_graph(12, 30000, 6, 16), 1e-9, 400)]      # then a smaller graph
    for (V, s, d, w, t), tol, max_iter in cases:
        rc1, r1, i1 = _forward(lib.cytom_ppr_weighted, V, s, d, w, t, tol, max_iter)
        rc2, r2, i2 = _forward(lib.cytom_ppr_weighted_noalloc, V, s, d, w, t, tol, max_iter)
        assert rc1 == 0 and rc2 == 0, (rc1, rc2, lib.cytom_ppr_last_error())
        assert i1 == i2 and i1 > 0
        assert np.array_equal(r1, r2), (V, len(s))
        assert np.isfinite(r2).all() and r2.sum() > 0


def test_noalloc_forward_refuses_graphs_beyond_its_scratch():
    lib = C.lib()
    V, E = 1 << 19, 1 << 20
    rng = np.random.default_rng(1)
    s = rng.integers(0, V, E).astype(np.uint32)
    d = rng.integers(0, V, E).astype(np.uint32)
    w = np.ones(E, np.float32)
    t = np.full(V, 1.0 / V, np.float32)
    rc, _, _ = _forward(lib.cytom_ppr_weighted_noalloc, V, s, d, w, t, 0.0, 10)
    assert rc == C.ERANGE
    assert b"allocation-free" in lib.cytom_ppr_last_error()
    # weights are still validated
    w[7] = 0.0
    rc, _, _ = _forward(lib.cytom_ppr_weighted_noalloc, 1 << 12, s % 4096, d % 4096, w,
                        np.full(4096, 1.0 / 4096, np.float32), 0.0, 10)
    assert rc == C.EINVAL
