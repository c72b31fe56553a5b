"""Build recipe of oracle/_ref/libref.so — TEST INFRASTRUCTURE.

    python oracle/ref_build.py [reference root, default /root/reference]

Compiles oracle/ref_driver.cpp (ours) against the reference's own standalone
headers sobol.h, saka.h and interval.h, included unmodified from the
reference tree (g++ -I <root>).  Nothing is copied: the headers are read
where they lie and only the built library lands in oracle/_ref/.  The rest of
the reference's path (pk.h, minimum_lbvh.h, main.cpp, the .cu kernels) needs
prlib, Orochi and Windows headers that are absent, so it is not built.

Where the reference tree is absent (the GPU host) build() does nothing and
returns None; the tests then fall back to the committed golden vectors.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
OUT = os.path.join(OUT_DIR, "libref.so")
DEFAULT_ROOT = "/root/reference"
HEADERS = ("sobol.h", "saka.h", "interval.h")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-w"]


def build(root: str = DEFAULT_ROOT, verbose: bool = False) -> str | None:
    if not all(os.path.isfile(os.path.join(root, h)) for h in HEADERS):
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = ["g++", *FLAGS, "-I", root, os.path.join(HERE, "ref_driver.cpp"), "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 and not sys.argv[1].startswith("-") else DEFAULT_ROOT,
                verbose="-v" in sys.argv))
