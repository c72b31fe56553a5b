"""Build recipe of the CPU oracle (test infrastructure): oracle/liboracle.so.

    python oracle/build.py

g++ -O3 with -ffp-contract=off and no fast-math, so the FP32 restatement does
the same IEEE operations in the same order as the gfx950 kernels.  The target
is x86-64-v3 (AVX2/FMA hosts) rather than -march=native so that the library
built here also loads on the GPU host.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["lbvh.cpp", "path.cpp", "interval.cpp"]
OUT = os.path.join(HERE, "liboracle.so")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-march=x86-64-v3", "-ffp-contract=off",
         "-fno-fast-math", "-Wall", "-Wextra", "-Wno-unused-parameter", "-pthread"]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    cmd = ["g++", *FLAGS, "-o", OUT + ".tmp", *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


SANITIZED = os.path.join(HERE, "_build", "liboracle_asan.so")


def build_sanitized(verbose: bool = False) -> str:
    """The same sources under AddressSanitizer and UndefinedBehaviorSanitizer
    (host code, -O1, frame pointers kept), for the sanitizer test; load it in
    a process that preloads libasan (see tests/test_sanitizers.py)."""
    os.makedirs(os.path.dirname(SANITIZED), exist_ok=True)
    flags = [f for f in FLAGS if f != "-O3"] + ["-O1", "-g", "-fno-omit-frame-pointer",
                                                  "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    cmd = ["g++", *flags, "-o", SANITIZED + ".tmp", *[os.path.join(HERE, s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(SANITIZED + ".tmp", SANITIZED)
    return SANITIZED


if __name__ == "__main__":
    print(build_sanitized(verbose=True) if "--sanitize" in sys.argv else build(verbose="-v" in sys.argv))
