"""Build recipe of libpalz.so (gfx950 HIP kernels + C ABI), in-tree.

    python -m palzilla_amd.build

Every source compiles for gfx950 only, with -ffp-contract=off (bit-exact FP32
against the CPU oracle) and packed FP32 ops disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libpalz.so")
SOURCES = [  # the longest compiles first: they run in parallel
    "pathcuts_k3.hip", "pathcuts_k3_records.hip", "pathcuts_k2.hip", "adjoint.hip", "batched.hip", "lbvh.hip",
    "pathcuts_k1.hip", "primary.hip", "solve.hip", "cull.hip", "trace.hip", "sort.hip", "prims.hip", "spectral.hip",
    "runtime.hip", "pathcuts.hip", "scene_io.hip", "renderer.hip", "sort_noalloc.hip", "rays.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-command-line-argument",
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
           "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    _build_host_program(verbose)
    return LIB


HOST_PROGRAM = os.path.join(PKG, "palz_render")


def _build_host_program(verbose: bool) -> None:
    """palz_render: the C++ host program over the renderer, linked against
    libpalz.so next to it (g++, host code only)."""
    tmp = HOST_PROGRAM + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "render_main.cpp"), "-o", tmp, "-L", PKG, "-lpalz", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"palz_render build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, HOST_PROGRAM)


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
