"""Build libmagace.so (HIP for gfx950) in-tree with hipcc.

Usage: python mag-ace_amd/build.py [--jobs N]
Objects go to mag-ace_amd/_build/, the library to mag-ace_amd/libmagace.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmagace.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MAGACE_ARCH", "gfx950")

SOURCES = ["device.hip", "neighbors.hip", "md.hip", "noalloc.hip", "api.cpp", "basis.cpp", "potential.cpp", "splines.cpp", "pair_host.cpp",
           "host_pool.cpp"]
# neighbour lists reproduce the reference's FP64 values bit for bit: no contraction
EXTRA = {"neighbors.hip": ["-ffp-contract=off"]}
# packed FP32 instructions are disabled on purpose (the compiler would otherwise form
# v_pk_* from ordinary float code)
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, src + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), p.stdout, p.stderr))
    warn = [l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if warn:
        sys.stderr.write("\n".join(warn[-40:]) + "\n")
    return obj


def build(jobs: int = 4) -> str:
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (p.stdout, p.stderr))
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs))
