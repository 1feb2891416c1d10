"""Builds tests/shim/execute_shim, a C++ program that includes
include/cytom.h and links the in-tree libcytom_hip.so (host compiler only).

Usage: python tests/shim/build_shim.py [output]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "execute_shim.cpp")
OUT = os.path.join(HERE, "execute_shim")


def build(out=OUT):
    lib_dir = os.path.join(ROOT, "cytom_amd")
    rel = os.path.relpath(lib_dir, os.path.dirname(os.path.abspath(out)))
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), SRC, "-o", out,
           "-L", lib_dir, "-lcytom_hip",
           "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("shim build failed:\n%s\n%s" % (r.stdout, r.stderr))
    return out


if __name__ == "__main__":
    print(build(*sys.argv[1:]))
