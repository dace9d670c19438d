"""Build libgeofem.so for gfx950 with hipcc (no CMake, no JIT cache).

    python -m geo_simulator_amd.csrc.build        # or build() from Python

Objects go to geo_simulator_amd/_build/, the library to
geo_simulator_amd/libgeofem.so (in-tree, so it ships with the source tree).
Element, assembly and ILU(0) kernels are compiled with -ffp-contract=off so
their results are bit-identical to the FP64 restatement of the reference order;
the solver keeps contraction (FMA) on.  Packed-FP32 instructions are disabled
for every GPU object.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
BUILD_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgeofem.so")
ARCH = os.environ.get("GEOFEM_ARCH", "gfx950")

SOURCES = [
    ("elements.hip", ["-ffp-contract=off"]),
    ("assembly.hip", ["-ffp-contract=off"]),
    ("operator.hip", []),
    ("solver.hip", []),
    ("measure.hip", []),
    ("ilu.hip", ["-ffp-contract=off"]),
    ("amg.hip", ["-ffp-contract=off"]),
    ("capi.cpp", []),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function", "-Wno-unused-result",
        "-I", os.path.join(ROOT, "include"), "-I", HERE,
    ]


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src, extra in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [cc, *common_flags(), *extra, "-x", "hip", "-c",
               os.path.join(HERE, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if out.strip() and (verbose or p.returncode != 0):
            print(f"--- {src}\n{out}", file=sys.stderr)
        if p.returncode != 0:
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
