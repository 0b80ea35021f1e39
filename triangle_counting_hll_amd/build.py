"""Builds ``libtchll.so`` (the HIP/gfx950 kernels and their C ABI) in-tree.

Usage: ``python -m triangle_counting_hll_amd.build`` (or ``build()`` from
``__graft_entry__``).  Each ``csrc/*.hip`` file is compiled with
``hipcc --offload-arch=gfx950`` and linked into one shared library next to
this file.  Packed FP32 instruction selection is disabled for all device code.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_NAME = "libtchll.so"
ARCH = os.environ.get("TCHLL_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-8000:]}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile every HIP source and link ``libtchll.so``; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    workers = max(1, min(len(srcs), (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(HERE, LIB_NAME)
    tmp = out + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-8000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(srcs)} HIP sources", flush=True)
    return out


def build_oracle(verbose: bool = True) -> str:
    """Compile the C oracle / CPU baseline (oracle/libtc_oracle.so)."""
    odir = os.path.join(ROOT, "oracle")
    res = subprocess.run(["make", "-C", odir, "-B", "libtc_oracle.so"], capture_output=True,
                         text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stderr[-4000:]}")
    path = os.path.join(odir, "libtc_oracle.so")
    if verbose:
        print(f"built {path}", flush=True)
    return path


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
