"""Build oracle/_ref/libref_rmat.so from the reference's own source text.

The reference's hot-path headers cannot be compiled here: util.h pulls in
schema_generated.h (flatbuffers) and tile-manager.h needs gtest, neither of
which exists in this image.  The two self-contained pieces the update
stream depends on are compiled instead, taken verbatim from the reference
tree at build time:

  * util::hash(uint32_t), util::hash(uint64_t), util::hashDouble
        src/include/util/util.h:178-200
  * RMATContext::rMat, RMATContext::rMatIterative and RMATContext::getEdge
        src/include/edge-inserter/rmat-context.h:19-73

The extracted function bodies are wrapped in a small harness whose only
additions are the thresholds the reference constructor never initialises
(defect D1; a = 0.57f, b = c = 0.19f from datatypes.h:17-19), the two
members the constructor does set (rmat-context.h:14-15: hash_ = hash(seed);
count_vertices_ is its argument) and an extern "C" entry point.  The generated source is deleted after compiling:
only the shared library stays under oracle/_ref/ (git-ignored).

Usage: python oracle/ref_build.py [REFERENCE_ROOT]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT_DIR, "libref_rmat.so")


def _extract(text, signature_regex):
    """Return the full text of the first function whose head matches."""
    m = re.search(signature_regex, text)
    if not m:
        raise RuntimeError("signature not found: " + signature_regex)
    start = m.start()
    i = text.index("{", m.end() - 1)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[start:j + 1]
    raise RuntimeError("unbalanced braces for " + signature_regex)


def build(ref_root="/root/reference", quiet=False):
    util_h = os.path.join(ref_root, "src/include/util/util.h")
    rmat_h = os.path.join(ref_root, "src/include/edge-inserter/rmat-context.h")
    if not (os.path.exists(util_h) and os.path.exists(rmat_h)):
        if not quiet:
            print("reference tree not present; oracle/_ref not built")
        return False
    util = open(util_h).read()
    rmat = open(rmat_h).read()
    h32 = _extract(util, r"inline\s+uint32_t\s+hash\s*\(\s*uint32_t\s+a\s*\)\s*\{")
    h64 = _extract(util, r"inline\s+uint64_t\s+hash\s*\(\s*uint64_t\s+a\s*\)\s*\{")
    hd = _extract(util, r"inline\s+double\s+hashDouble\s*\(\s*uint64_t\s+i\s*\)\s*\{")
    rec = _extract(rmat, r"edge_t\s+rMat\s*\(\s*uint64_t\s+n\s*,[^)]*\)\s*\{")
    it = _extract(rmat, r"edge_t\s+rMatIterative\s*\(\s*uint64_t\s+n\s*,[^)]*\)\s*\{")
    get = _extract(rmat, r"edge_t\s+getEdge\s*\(\s*uint64_t\s+index\s*\)\s*\{")
    src = "\n".join([
        "#include <cstdint>",
        "namespace util {",
        h32, h64, hd,
        "}",
        "struct edge_t { uint64_t src; uint64_t tgt; };",
        "struct Harness {",
        "  double a_, ab_, abc_;",
        "  uint64_t count_vertices_, hash_;",
        rec,
        it,
        get,
        "};",
        'extern "C" uint64_t ref_hash64(uint64_t a) { return util::hash(a); }',
        'extern "C" uint32_t ref_hash32(uint32_t a) { return util::hash(a); }',
        'extern "C" double ref_hash_double(uint64_t a) { return util::hashDouble(a); }',
        'extern "C" void ref_rmat(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,',
        '                         int recursive, uint32_t* s, uint32_t* d) {',
        "  Harness h;",
        "  h.a_ = 0.57f; h.ab_ = h.a_ + 0.19f; h.abc_ = h.ab_ + 0.19f;",
        "  h.count_vertices_ = 1ull << scale;",
        "  h.hash_ = util::hash(seed);",
        "  for (uint64_t j = 0; j < n; ++j) {",
        "    const uint64_t index = first + j;",
        "    edge_t e;",
        "    if (recursive) {",
        "      e = h.rMat(h.count_vertices_, util::hash(2ul * index * h.hash_),",
        "                 util::hash((2ul * index + 1) * h.hash_));",
        "    } else {",
        "      e = h.getEdge(index);",
        "    }",
        "    s[j] = (uint32_t)e.src; d[j] = (uint32_t)e.tgt;",
        "  }",
        "}",
    ])
    os.makedirs(OUT_DIR, exist_ok=True)
    gen = os.path.join(OUT_DIR, "ref_rmat_gen.cpp")
    with open(gen, "w") as f:
        f.write(src)
    try:
        subprocess.check_call(["g++", "-O2", "-std=c++14", "-fPIC", "-shared",
                               "-ffp-contract=off", "-o", LIB, gen])
    finally:
        os.remove(gen)
    if not quiet:
        print("built", LIB)
    return True


if __name__ == "__main__":
    build(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
