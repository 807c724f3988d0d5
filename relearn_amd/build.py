"""Builds librelearn_bh.so (the HIP kernels and the C ABI of include/relearn_bh.h) for gfx950.

Each csrc/*.hip file is compiled by hipcc into an object, then the objects are linked into
relearn_amd/librelearn_bh.so.  FP contraction is off for every translation unit (the FP64 path must
round like the CPU algorithm) and packed FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librelearn_bh.so")

SOURCES = ["rl_api.hip", "rl_tree.hip", "rl_scan.hip", "rl_search.hip", "rl_search_coop.hip", "rl_search_walk.hip",
           "rl_commit.hip",
           "rl_step.hip", "rl_net.hip"]

ARCH = os.environ.get("RELEARN_OFFLOAD_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, flush=True)
    return obj


HOST_DIR = os.path.join(HERE, "host")
HOST_EXAMPLE = os.path.join(HOST_DIR, "update_example")


def build_host_example(verbose: bool = False) -> str:
    """Compiles the C++ adapter example (host code only, g++) against librelearn_bh.so."""
    cxx = os.environ.get("CXX", shutil.which("g++") or "g++")
    tmp = HOST_EXAMPLE + ".tmp"
    cmd = [cxx, "-std=c++20", "-O2", "-Wall", "-Wextra", os.path.join(HOST_DIR, "update_example.cpp"), "-o", tmp,
           f"-L{HERE}", "-lrelearn_bh", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host example build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, HOST_EXAMPLE)
    return HOST_EXAMPLE


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_host_example(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
