"""Algorithm store: named vertex programs for the launcher.

The reference compiles each added algorithm source to LLVM bitcode and links
the bitcode of a task into one executable (scripts/run.py:683-759
AlgorithmStore, examples/*.cc main_<name>).  Here every vertex program is a
device program of libcytom_hip.so, so the store maps a name to one of them:
``--add --name N --path P`` reads the source file P and binds N to the device
program whose entry point P defines (main_bfs, main_cc, main_sssp,
main_pagerank_delta, main_ppr).  P may also name a built-in program
directly.  The registry is a JSON file (CYTOM_ALGORITHM_STORE, default
``algorithm-store.json`` beside this module), as the reference keeps
ALGORITHMS in scripts/config.py.
"""
import json
import os
import re

from . import BFS, CC, PPR, PRD, SSSP

PREFIX = "-----------[AlgorithmStore]: "

# Program names the engine accepts in a job's --algorithm.  The -async
# variants of the reference (examples/bfs-async.cc, cc-async.cc) reach the
# same fixpoint as their synchronous forms, which is what the device programs
# compute.
BUILTIN = {
    "bfs": BFS, "bfs-async": BFS,
    "cc": CC, "cc-async": CC, "connected-components": CC, "connected-components-async": CC,
    "sssp": SSSP, "sssp-async": SSSP,
    "pagerank-delta": PRD, "pagerankdelta": PRD,
    "ppr": PPR,
}
KIND_NAMES = {BFS: "bfs", CC: "cc", SSSP: "sssp", PRD: "pagerank-delta", PPR: "ppr"}
_ENTRY = re.compile(r"\bmain_(bfs|cc|sssp|pagerank_delta|ppr)\s*\(")
_ENTRY_KIND = {"bfs": BFS, "cc": CC, "sssp": SSSP, "pagerank_delta": PRD, "ppr": PPR}


def default_path():
    return os.environ.get("CYTOM_ALGORITHM_STORE",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                       "algorithm-store.json"))


def program_of_source(path):
    """Device program kind for an algorithm source file or built-in name."""
    if path.lower() in BUILTIN:
        return BUILTIN[path.lower()]
    try:
        with open(path, errors="replace") as f:
            text = f.read()
    except OSError:
        return None
    kinds = {_ENTRY_KIND[m.group(1)] for m in _ENTRY.finditer(text)}
    return kinds.pop() if len(kinds) == 1 else None


class AlgorithmStore:
    def __init__(self, path=None, out=print):
        self.path = path or default_path()
        self.out = out
        self.algorithms = {}  # name -> {"kind": int, "path": str}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.algorithms = json.load(f)

    def save(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.algorithms, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def resolve(self, name):
        """Program kind of a job's --algorithm: store entries first, then the
        built-in names; None when neither knows it."""
        if name in self.algorithms:
            return self.algorithms[name]["kind"]
        return BUILTIN.get(name.lower())

    # run.py:691-759 actions
    def connect(self, args):
        if args.add:
            self.add(args.name, args.path)
        elif args.delete:
            self.delete(args.name)
        elif args.clear:
            self.clear()
        elif args.list:
            self.out("\n".join(sorted(self.algorithms)))
        else:
            self.out(PREFIX + "Please specify action (add/delete/list)")

    def add(self, names, paths):
        if not names:
            self.out(PREFIX + "Please specify name(s) of the algorithm(s)")
            return False
        if not paths:
            self.out(PREFIX + "Please specify path(s) of the algorithm(s)")
            return False
        if len(names) != len(paths):
            self.out(PREFIX + "Oops! The number of names does not equal to the number of paths.")
            return False
        ok = True
        for name, path in zip(names, paths):
            kind = program_of_source(path)
            if kind is None:
                self.out(PREFIX + "%s defines no vertex program with a device implementation "
                         "(one of main_bfs, main_cc, main_sssp, main_pagerank_delta, main_ppr)"
                         % path)
                ok = False
                continue
            existed = name in self.algorithms
            self.algorithms[name] = {"kind": kind, "path": os.path.abspath(path)
                                     if os.path.exists(path) else path}
            self.out(PREFIX + ("Successfully overwrite " if existed else "Successfully add ")
                     + name)
        self.save()
        return ok

    def delete(self, names):
        if not names:
            self.out(PREFIX + "Please specify name(s) of the algorithm(s)")
            return False
        gone = []
        for name in names:
            if name not in self.algorithms:
                self.out(PREFIX + "Oops! %s is not a valid algorithm name" % name)
            else:
                gone.append(name)
                del self.algorithms[name]
        self.save()
        self.out(PREFIX + "Successfully delete " + " ".join(gone))
        return True

    def clear(self):
        self.algorithms = {}
        self.save()
        self.out(PREFIX + "Successfully delete all algorithms")
        return True
