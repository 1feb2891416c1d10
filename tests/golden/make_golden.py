"""Write tests/golden/rmat_kat.json from the reference's own hash / RMAT code.

The vectors come from oracle/_ref/libref_rmat.so, which oracle/ref_build.py
compiles from the reference tree's util.h and rmat-context.h text
(hash, hashDouble, rMat, rMatIterative, getEdge).  They pin the oracle's
restatement of the stream generator on machines without the reference tree.

Usage: python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref_build  # noqa: E402


def main():
    if not os.path.exists(ref_build.LIB):
        if not ref_build.build():
            raise SystemExit("reference tree absent: cannot regenerate the fixtures")
    R = ctypes.CDLL(ref_build.LIB)
    R.ref_hash64.argtypes = [ctypes.c_uint64]
    R.ref_hash64.restype = ctypes.c_uint64
    R.ref_hash32.argtypes = [ctypes.c_uint32]
    R.ref_hash32.restype = ctypes.c_uint32
    R.ref_hash_double.argtypes = [ctypes.c_uint64]
    R.ref_hash_double.restype = ctypes.c_double
    u32p = np.ctypeslib.ndpointer(np.uint32)
    R.ref_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                           ctypes.c_int, u32p, u32p]
    rng = np.random.default_rng(20261203)
    hx = [0, 1, 2, 3, 0xFFFFFFFFFFFFFFFF] + [int(x) for x in rng.integers(0, 2**63, 27,
                                                                          dtype=np.uint64)]
    h32x = [0, 1, 0xFFFFFFFF] + [int(x) for x in rng.integers(0, 2**32, 13, dtype=np.uint64)]
    out = {
        "source": "reference util.h:178-200 and rmat-context.h:19-73 via oracle/_ref",
        "hash64": [[x, int(R.ref_hash64(x))] for x in hx],
        "hash32": [[x, int(R.ref_hash32(x))] for x in h32x],
        "hash_double": [[x, float(R.ref_hash_double(x)).hex()] for x in hx],
        "rmat": [],
    }
    for scale, seed, first, n in [(1, 2, 0, 8), (5, 2, 0, 32), (10, 7, 1000, 32),
                                  (16, 2, 0, 64), (16, 2, 1048000, 32), (22, 2, 0, 64),
                                  (22, 2, 67108800, 32), (27, 2, 2147483600, 32)]:
        s = np.empty(n, np.uint32)
        d = np.empty(n, np.uint32)
        R.ref_rmat(scale, seed, first, n, 0, s, d)
        out["rmat"].append({"scale": scale, "seed": seed, "first": first,
                            "src": s.tolist(), "dst": d.tolist()})
    with open(os.path.join(HERE, "rmat_kat.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "rmat_kat.json"))


if __name__ == "__main__":
    main()
