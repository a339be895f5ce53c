"""Recipe: build reference routines with their module closures into
oracle/_ref with AMD flang, compiling each reference file where it lies
under /root/reference (never copied): `python3 ref_closure.py branch`
(split_join, qmc/dmc/dmc_elec/split_join_mod.f90 -> _ref/libref_branch.so)
or `python3 ref_closure.py ylm` (the recursive real spherical harmonics of
qmc/vmc/basis_fns.f90 -> _ref/libref_ylm.so, with ref_ylm_shim.f90).

TEST INFRASTRUCTURE (the oracle's checker), run by `make ref`.  The module
closure of split_join_mod is found from the `use` statements (118 files),
compiled in dependency order, and linked with the reference's rannyu.f90,
lib/lib/shell.f (the sort split_join calls) and ref_branch_shim.f90 (ours:
bind(C) entry points).  Module procedures in the closure call 24 external
routines (hpsi, tmove, LAPACK, ...) that split_join never reaches; they stay
unresolved and the library is loaded with lazy binding.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("REF", "/root/reference/qmc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
# target: (build folder, library, root files of the closure, extra sources)
TARGETS = {
    "branch": ("branch", "libref_branch.so", ["dmc/dmc_elec/split_join_mod.f90"],
               [(os.path.join(REF, "vmc/rannyu.f90"), "free"),
                (os.path.join(os.path.dirname(REF), "lib/lib/shell.f"), "fixed"),
                (os.path.join(HERE, "ref_branch_shim.f90"), "free")]),
    # the recursive real spherical harmonics of qmc/vmc/basis_fns.f90 (its
    # module closure; basis_fns's own routines are never called)
    "ylm": ("ylm", "libref_ylm.so", ["vmc/basis_fns.f90"],
            [(os.path.join(HERE, "ref_ylm_shim.f90"), "free")]),
}
OUT = SO = ROOTS = EXTRA = None


def _select(name):
    global OUT, SO, ROOTS, EXTRA
    d, so, roots, extra = TARGETS[name]
    OUT = os.path.join(HERE, "_ref", d)
    SO = os.path.join(HERE, "_ref", so)
    ROOTS, EXTRA = roots, extra


def _sources():
    txt, mod2file = {}, {}
    for f in glob.glob(REF + "/**/*.f90", recursive=True) + glob.glob(REF + "/**/*.f", recursive=True):
        rel = os.path.relpath(f, REF)
        t = open(f, errors="replace").read()
        txt[rel] = re.sub(r"!.*", "", t)
        for m in re.findall(r"^\s*module\s+(\w+)\s*$", txt[rel], flags=re.M | re.I):
            mod2file.setdefault(m.lower(), rel)
    return txt, mod2file


def _levels(roots):
    """Files of the closure grouped so that each group only uses modules of
    earlier groups."""
    txt, mod2file = _sources()
    deps = {}
    stack = list(roots)
    while stack:
        f = stack.pop()
        if f in deps:
            continue
        used = {m.lower() for m in re.findall(r"^\s*use\s+(\w+)", txt[f], flags=re.M | re.I)}
        deps[f] = {mod2file[m] for m in used if m in mod2file} - {f}
        stack.extend(deps[f])
    level = {}

    def lev(f, seen=()):
        if f not in level:
            level[f] = 1 + max([lev(g, seen + (f,)) for g in deps[f] if g not in seen] or [-1])
        return level[f]

    for f in deps:
        lev(f)
    out = [[] for _ in range(max(level.values()) + 1)]
    for f, l in level.items():
        out[l].append(f)
    return out


def _compile(src, form):
    obj = os.path.join(OUT, os.path.relpath(src, "/").replace("/", "__") + ".o")
    flags = ["-ffree-form"] if form == "free" else ["-ffixed-form", "-ffixed-line-length=132"]
    cmd = [FLANG, "-cpp", *flags, "-fPIC", "-O2", "-w", "-module-dir", os.path.join(OUT, "mod"),
           "-I", os.path.join(OUT, "mod"), "-I", os.path.dirname(src), "-c", "-o", obj, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"flang failed on {src}:\n{r.stderr[-2000:]}")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(os.path.join(OUT, "mod"), exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for group in _levels(ROOTS):
            srcs = [(os.path.join(REF, f), "free" if f.endswith(".f90") else "fixed") for f in group]
            objs += list(ex.map(lambda a: _compile(*a), srcs))
        objs += [_compile(src, form) for src, form in EXTRA]
    r = subprocess.run([FLANG, "-shared", "-fPIC", "-o", SO, *objs], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stderr[-2000:])
    return SO


if __name__ == "__main__":
    target = sys.argv[1] if len(sys.argv) > 1 else "branch"
    _select(target)
    if not os.path.isdir(REF):
        print(f"reference sources absent: {os.path.basename(SO)} not built")
        sys.exit(0)
    print(build())
