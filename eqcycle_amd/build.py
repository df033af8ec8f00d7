"""Build libeqcycle_hip.so for gfx950 in-tree (eqcycle_amd/libeqcycle_hip.so).

    python -m eqcycle_amd.build

HIP sources are compiled by hipcc for gfx950 only; the host-only AMG set-up
is compiled by g++.  Packed FP32 instructions are disabled
(-target-feature -packed-fp32-ops) for every device object.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libeqcycle_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-I", CSRC, "-I", INCLUDE, "-Wno-unused-result"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-I", CSRC, "-I", INCLUDE, "-I", "/opt/rocm/include",
             "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["eq_kernels.hip", "eq_api.hip", "eq_learn.hip"]
CXX_SOURCES = ["amg_host.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False, defines=(), out: str | None = None) -> str:
    """defines: extra preprocessor definitions, e.g. ("EQ_TILE_I=32",) for
    32 x 16 tiles in the grid kernels (default 64 x 16); out: where to write
    the library (default: the package's own, which the package loads)."""
    lib = out or LIB
    obj_dir = OUT_DIR if out is None else os.path.join(os.path.dirname(os.path.abspath(out)), "_build")
    os.makedirs(obj_dir, exist_ok=True)
    dflags = [f"-D{d}" for d in defines]
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, *dflags, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(obj_dir, src.replace(".cpp", ".o"))
        jobs.append(["g++", *CXX_FLAGS, *dflags, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=4) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--define", action="append", default=[], help="preprocessor definition NAME=VALUE")
    ap.add_argument("--out", default=None, help="library path (default: the package's)")
    a = ap.parse_args()
    print(build(verbose=True, defines=a.define, out=a.out))
