"""Build recipe of libchamp_amd.so (gfx950).

hipcc compiles the C ABI and its kernels into one in-tree shared library.
Packed FP32 instructions are switched off (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libchamp_amd.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ORDERS = (2, 3, 4, 5, 6, 7)      # e-e-n polynomial orders, one object each

FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose: bool = False, jobs: int | None = None, out: str | None = None,
          build_dir: str | None = None) -> str:
    """Compile engine_api.hip and one engine_ord.hip object per e-e-n order
    in parallel, then link them into libchamp_amd.so (or `out`)."""
    OUT = out or globals()["OUT"]
    BUILD = build_dir or globals()["BUILD"]
    os.makedirs(BUILD, exist_ok=True)
    units = [(os.path.join(CSRC, "engine_api.hip"), os.path.join(BUILD, "engine_api.o"), [])]
    for o in ORDERS:
        units.append((os.path.join(CSRC, "engine_ord.hip"), os.path.join(BUILD, f"engine_ord{o}.o"),
                      [f"-DQMCW_ORD={o}"]))
    jobs = jobs or min(len(units), max(1, os.cpu_count() or 1))
    pending = list(units)
    running = []
    errors = []
    while pending or running:
        while pending and len(running) < jobs:
            src, obj, extra = pending.pop(0)
            cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", obj, src]
            running.append((obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                                  text=True)))
        obj, p = running.pop(0)
        out, err = p.communicate()
        if p.returncode != 0:
            errors.append(f"{obj}:\n{err[-3000:]}")
        elif verbose and err:
            print(err[-1000:])
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp",
           *[obj for _, obj, _ in units]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
