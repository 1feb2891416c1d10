"""C-ABI surface of libcytom_hip.so (no GPU needed): every symbol that
include/cytom.h declares is exported and bound, and the host-side stream
sampler agrees with the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

import cytom_amd as C
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cytom.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(cytom_[a-z_0-9]+)\s*\(", text,
                       flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    assert "cytom_create" in names and "cytom_update_batch_device" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(C.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_python_bindings_cover_header():
    assert sorted(C.SIGNATURES) == declared_functions()
    C.lib()  # binds every signature


def test_struct_sizes_match_header_layout():
    assert ctypes.sizeof(C.Config) == 8 + 4 + 4 + 8 + 8 + 4 + 4
    assert ctypes.sizeof(C.RunStats) == 4 * 4 + 8 * 3 + 8
    assert ctypes.sizeof(C.BatchStats) == 8 * 4 + 4 * 4 + 8
    assert ctypes.sizeof(C.Tuning) == 4 * 7


def test_invalid_arguments_fail_without_a_device():
    L = C.lib()
    h = ctypes.c_void_p()
    assert L.cytom_create(None, ctypes.byref(h)) == C.EINVAL
    cfg = C.Config(0, 0, 0, 0.0, 0, 0, 0)
    assert L.cytom_create(ctypes.byref(cfg), ctypes.byref(h)) == C.EINVAL
    cfg = C.Config(1 << 10, 100, 0, 0.0, 0, 0, 0)  # tile width not a power of two
    assert L.cytom_create(ctypes.byref(cfg), ctypes.byref(h)) == C.EINVAL
    assert L.cytom_destroy(None) == C.EINVAL
    assert L.cytom_execute(None, 0, None) == C.EINVAL


def test_stream_sampler_matches_oracle():
    a = C.Stream(17)
    b = O.Stream(17)
    for n_ins, n_del in [(1000, 0), (500, 120), (0, 33), (2000, 700), (10, 5000)]:
        fa, da = a.next(n_ins, n_del)
        fb, db = b.next(n_ins, n_del)
        assert fa == fb
        assert np.array_equal(da, db)


def test_update_arguments_checked_before_the_library_is_called():
    # A Graph shell with no context: every case below must be rejected by the
    # binding itself (the library would read n*4 bytes from each array).
    g = C.Graph.__new__(C.Graph)
    g._h = None
    u = lambda *v: np.array(v, np.uint32)
    with pytest.raises(ValueError, match="lengths differ"):
        g.update(u(1, 2, 3), u(4, 5))
    with pytest.raises(ValueError, match="lengths differ"):
        g.update(u(1, 2), u(4, 5), u(1))
    with pytest.raises(ValueError, match="lengths differ"):
        g.update(u(), u(), None, u(1, 2), u(3), None)
    with pytest.raises(ValueError, match="lengths differ"):
        g.update(u(), u(), None, u(1), u(3), u(1, 1))
    with pytest.raises(ValueError, match="non-integral"):
        g.update(u(1), u(2), np.array([1.5], np.float32))
    with pytest.raises(ValueError, match="zero weights"):
        g.update(u(1), u(2), np.array([0.0]))
    with pytest.raises(ValueError, match="outside"):
        g.update(np.array([-1]), u(2))
    with pytest.raises(ValueError, match="outside"):
        g.update(u(1), u(2), np.array([1 << 33]))
    # integral floats are accepted as ids/weights
    assert C._weights(np.array([3.0, 255.0]), "w").tolist() == [3, 255]
