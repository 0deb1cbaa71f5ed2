"""Build recipe for the native libraries of ugsdr_amd (run in-tree).

  libugsdr_hip.so    HIP kernels + C ABI (include/ugsdr_hip.h), gfx950 only
  libugsdr_siggen.so deterministic synthetic GNSS signal generator (host C)

Usage: python -m ugsdr_amd.build  (or ugsdr_amd.build.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(ROOT, "_build", "obj")
HIP_LIB = os.path.join(HERE, "libugsdr_hip.so")
GEN_LIB = os.path.join(HERE, "libugsdr_siggen.so")

ARCH = os.environ.get("UGSDR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed FP32 VALU instructions are disabled for every device compilation.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++20", "-Wall", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]

HIP_SOURCES = ["acq.hip", "corr1023.hip", "ops.hip", "bank.hip", "reduced.hip", "server.hip", "dfe.hip", "tracker.hip", "jse.hip", "noalloc.hip", "lib.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {os.path.basename(cmd[-1])}")
    return r


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", "-o", obj, path])
    return obj


def build(jobs: int | None = None) -> None:
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    tmp = GEN_LIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-pthread", "-o", tmp,
          os.path.join(CSRC, "siggen.c"), "-lm"])
    os.replace(tmp, GEN_LIB)


if __name__ == "__main__":
    build()
    print(HIP_LIB)
