#!/usr/bin/env python3
"""Benchmark of the MI355X caustic hot path; prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lbvh]

Workload `lbvh`: the LBVH + onesweep build of 16,777,216 triangles (centroids
U[0,1]^3, lognormal sizes, random orientation; 30-bit Morton keys).  A step is
one full build (scene AABB, Morton keys + digit counts, four sort passes,
deltas, bottom-up build) over triangles already resident in HBM.

Only one GPU is used: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged specular paths/s at 1e-5 vertex err, batched vs plain; LBVH+sort Mtri/s"
HBM_PEAK_GBPS = 8000.0

# Algorithmic bytes per triangle of each LBVH stage (the reference's pass
# structure): AABB read 36; Morton read 36 + write 8; sort pass read 8 + write
# 8; deltas read 8 + write 1; build reads triangle 36 + index 4 + delta 1 and
# writes one 64-byte node per internal node.
STAGE_BYTES_PER_TRI = {
    "scene_aabb": 36.0, "morton_hist": 44.0, "sort_pass0": 16.0, "sort_pass1": 16.0,
    "sort_pass2": 16.0, "sort_pass3": 16.0, "deltas": 9.0, "build": 41.0 + 64.0,
}
LBVH_BYTES_PER_TRI = 266.0  # 202 B/tri + 64 B per internal node


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measure_copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> float:
    from palzilla_amd._lib import call
    from palzilla_amd.device import DeviceBuffer, Event, Stream
    import ctypes as C
    src, dst = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    src.zero()
    st = Stream()
    e0, e1 = Event(), Event()
    call("palz_copy_bytes", C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(nbytes), C.c_void_p(st.handle))
    best = 1e30
    for _ in range(reps):
        e0.record(st)
        call("palz_copy_bytes", C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(nbytes),
             C.c_void_p(st.handle))
        e1.record(st)
        best = min(best, e0.elapsed_ms(e1))
    src.free()
    dst.free()
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def bench_lbvh(args) -> dict:
    import numpy as np

    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Event, Stream, sync
    from palzilla_amd.lbvh import BVHGPUBuilder, OnesweepSort

    n = args.n
    log(f"[bench] generating {n} triangles")
    tris = scenes.triangle_soup(n, seed=2024)
    d_tris = DeviceBuffer.from_array(tris)
    st = Stream()
    sorter = OnesweepSort()
    sorter.reserve(n)
    b = BVHGPUBuilder()
    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
    st.sync()
    sync()
    e0, e1 = Event(), Event()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(args.steps):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
    e1.record(st)
    st.sync()
    sync()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1) / args.steps

    # per-kernel durations, live, on the launch stream
    b.set_timing(True)
    stage_ms = {k: 0.0 for k in b.STAGES}
    for _ in range(args.steps):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
        for k, v in b.timings().items():
            stage_ms[k] += v / args.steps
    b.set_timing(False)
    sort_pass_ms = sum(stage_ms[f"sort_pass{i}"] for i in range(4)) / 4.0
    kernel_total = {"k_sort_pass": 4 * sort_pass_ms, "k_build": stage_ms["build"],
                    "k_morton_hist": stage_ms["morton_hist"], "k_scene_aabb": stage_ms["scene_aabb"],
                    "k_deltas": stage_ms["deltas"]}
    dominant = max(kernel_total, key=kernel_total.get)
    per_launch_ms = {"k_sort_pass": sort_pass_ms, "k_build": stage_ms["build"],
                     "k_morton_hist": stage_ms["morton_hist"], "k_scene_aabb": stage_ms["scene_aabb"],
                     "k_deltas": stage_ms["deltas"]}[dominant]
    bytes_per_tri = {"k_sort_pass": 16.0, "k_build": STAGE_BYTES_PER_TRI["build"],
                     "k_morton_hist": 44.0, "k_scene_aabb": 36.0, "k_deltas": 9.0}[dominant]
    achieved = bytes_per_tri * n / (per_launch_ms * 1e-3) / 1e9

    # parity spot check of the timed configuration: sortedness and leaf count
    words, deltas = b.download_sorted()
    keys = (words & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.all(np.diff(keys) >= 0), "sorted keys are not sorted"
    assert np.array_equal(np.sort(words >> np.uint64(32)), np.arange(n, dtype=np.uint64))

    copy_gbps = measure_copy_bandwidth()
    mtri = n / (ms * 1e-3) / 1e6
    out = {
        "metric": METRIC,
        "value": round(mtri, 2),
        "unit": "Mtri/s",
        "ms_per_step": round(ms, 4),
        "dtype": "u32",
        "config": {"workload": "lbvh+onesweep, 16M-triangle soup (30-bit Morton keys)",
                   "n_triangles": n, "morton_bits": 30},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": None},
        "lbvh": {"stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
                 "algorithmic_bytes_per_tri": LBVH_BYTES_PER_TRI,
                 "achieved_GBps": round(LBVH_BYTES_PER_TRI * n / (ms * 1e-3) / 1e9, 1),
                 "hbm_frac": round(LBVH_BYTES_PER_TRI * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                 "copy_kernel_GBps": round(copy_gbps, 1)},
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu:
        import oracle
        sample = min(n, args.cpu_sample)
        log(f"[bench] cpu baseline on {sample} triangles, 1 thread")
        t0 = time.perf_counter()
        oracle.lbvh(tris[:sample], bits=30)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(sample / dt / 1e6, 3), "unit": "Mtri/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle LBVH build (scene AABB, Morton, stable sort, deltas, "
                                         f"serial climb) of the first {sample} triangles, 1 thread"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="lbvh", choices=["lbvh"])
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--cpu-sample", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); nothing is run across GPUs.",
              file=sys.stderr)
        return 2
    out = bench_lbvh(args)
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
