"""Builds libdiablo_gpu.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m diablo_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdiablo_gpu.so")
SOURCES = ["dgpu_kernels.hip", "dgpu_hot_x.hip", "dgpu_hot_z.hip", "dgpu_columns.hip", "dgpu_colseg.hip", "dgpu_stats.hip", "dgpu_les.hip", "dgpu_zdft.hip", "dgpu_poisson.hip",
           "dgpu_capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DGPU_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    # the transform stage loops ask to be unrolled, which only the builds for
    # a compile-time plan can do; the run-time plan builds keep the loop
    "-Wno-pass-failed",
]


def build(verbose: bool = False) -> str:
    """Compiles the sources in parallel (one hipcc per file), then links."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found; the HIP extension cannot be built")
    inc = ["-I", os.path.join(HERE, "..", "include")]
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with tempfile.TemporaryDirectory(prefix="dgpu_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]

        def compile_one(i):
            cmd = [HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, SOURCES[i]), "-o", objs[i]]
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        with ThreadPoolExecutor(jobs) as ex:
            results = list(ex.map(compile_one, range(len(SOURCES))))
        for src, r in zip(SOURCES, results):
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n" + r.stdout[-4000:] + r.stderr[-8000:])
            if verbose and r.stderr.strip():
                print(r.stderr[-2000:], file=sys.stderr)
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout[-4000:] + r.stderr[-8000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
