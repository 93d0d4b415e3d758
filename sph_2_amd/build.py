"""Build libsphx.so (gfx950) in-tree with hipcc.

Usage: python -m sph_2_amd.build [--jobs N]

Every HIP source is compiled with packed FP32 disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops) and the library is linked
into sph_2_amd/libsphx.so, which the ctypes loader picks up.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsphx.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = [
    "sphx_api.hip",
    "sphx_sort.hip",
    "sphx_hydro.hip",
    "sphx_gravity.hip",
    "sphx_plain.hip",
    "sphx_integrate.hip",
    "sphx_batch.hip",
    "sphx_bh.hip",
    "sphx_io.hip",
]

ARCH = os.environ.get("SPHX_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libsphx)")


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-variable",
        "-Wno-unused-but-set-variable",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        lines = r.stderr.splitlines()
        errs = [k for k, ln in enumerate(lines) if "error:" in ln]
        shown = "\n".join("\n".join(lines[k:k + 4]) for k in errs[:8]) or r.stderr[-6000:]
        raise RuntimeError(f"hipcc failed on {src}:\n{shown}")
    return obj


def build(jobs: int = 6, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
