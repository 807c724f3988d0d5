"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/bh_oracle.cpp`` into ``oracle/libbh_oracle.so`` with g++ -O3 -fopenmp and
-ffp-contract=off, without -march (the reference build sets none).

The reference's own sources cannot be compiled here: the Barnes-Hut headers include fmt, spdlog,
range-v3 and Boost, which the reference fetches from the network at configure time.  So there is no
``oracle/_ref`` build; the restatement is pinned by the checks listed in bh_oracle.cpp's header.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "bh_oracle.cpp")
LIB = os.path.join(HERE, "libbh_oracle.so")

CXXFLAGS = [
    "-std=c++20",
    "-O3",
    "-fopenmp",
    "-ffp-contract=off",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wextra",
]


def build(verbose: bool = False) -> str:
    cmd = ["g++", *CXXFLAGS, SRC, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
