#!/usr/bin/env python3
"""Benchmark of the MI355X caustic hot path; prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c4|lbvh]

Default workload `c2`, the headline: spectral two-interface refraction through
a glass slab (256x256-cell wavy top, 131,076 triangles, 2^22 photons, 1024^2
receiver, 16 wavelengths, Cauchy n = 1.5 + 4200 nm^2 / lambda^2).  Photons are
traced, the tuple table built and the (receiver point, tuple) list flattened
once; a step is then one batched solve of every (path, wavelength) item with
chain length 1..3: persistent FP32 Newton kernels, then throughput, dA/dw and
irradiance with one FP64 refinement step on every converged chain, per-item
results written to HBM.  The value is converged, contributing specular paths
per second (scaled by the fraction of a sample whose refined vertices lie
within 1e-5 of the FP64 oracle).  Beside it: the plain per-descriptor solver
(one launch per event descriptor, as the reference renders) on the same tuple
set, the CPU oracle solve of a bounded sample, and the sub-results `c4` (256
scenes in one hipGraph launch vs per-scene plain launches) and `lbvh` (16M
triangle LBVH + onesweep build) unless --no-extras.

Only one GPU is used: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "converged specular paths/s at 1e-5 vertex err, batched vs plain; LBVH+sort Mtri/s"
HBM_PEAK_GBPS = 8000.0
FP32_PEAK_TFLOPS = 157.3
CPU_THREADS = 16  # host cores per GPU on the benchmark machine

# Algorithmic bytes per triangle of each LBVH stage (the reference's pass
# structure): AABB read 36; Morton read 36 + write 8; sort pass read 8 + write
# 8; deltas read 8 + write 1; build reads triangle 36 + index 4 + delta 1 and
# writes one 64-byte node per internal node.
STAGE_BYTES_PER_TRI = {
    "scene_aabb": 36.0, "morton_hist": 44.0, "sort_pass0": 16.0, "sort_pass1": 16.0,
    "sort_pass2": 16.0, "sort_pass3": 16.0,
    "build": 8.0 + 36.0 + 1.0 + 64.0,  # sorted word + triangle + delta + internal node
    "build_levels": 0.0,  # the levels above the leaves: a few per cent of the nodes
}
LBVH_BYTES_PER_TRI = sum(STAGE_BYTES_PER_TRI.values())  # 253 B/tri
# The same build priced by the reference's separate passes (a count pass reads
# the keys, the deltas are a pass of their own, the build reads a 4-byte index):
# 36 + 44 + 8 + 4 x 16 + 9 + 41 per triangle plus 64 per internal node.
LBVH_REF_PASSES_BYTES = (36 + 44 + 8 + 4 * 16 + 9 + 41, 64)
LBVH_KERNEL = {"scene_aabb": "k_scene_aabb", "morton_hist": "k_morton_hist", "sort_pass0": "k_sort_pass<false>",
               "sort_pass1": "k_sort_pass<false>", "sort_pass2": "k_sort_pass<false>",
               "sort_pass3": "k_sort_pass<false>", "build": "k_build",
               "build_levels": "k_build_level"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measure_copy_bandwidth(nbytes: int = 1 << 30, reps: int = 5) -> float:
    from palzilla_amd._lib import call
    from palzilla_amd.device import DeviceBuffer, Event, Stream
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


# ---------------------------------------------------------------------------------------------------
# spectral slab (headline)

def NEWTON_BYTES(K: int) -> tuple[int, int]:
    """Compulsory HBM bytes of k_newton<K> as (per path, per item).  Per path
    (its L wavelength items sit next to each other and share these): the path
    record (64), its K triangles (36 each) and attributes (40 each) and the
    receiver point (40).  Per (path, wavelength) item: the wavelength (4) and,
    written, the converged flag (4) and either the solution (8K) and iteration
    count (4) or a 48-byte result record (the larger, 48, is counted)."""
    return 64 + 76 * K + 40, 4 + 4 + 48


def OCCLUDE_BYTES(K: int) -> tuple[int, int]:
    """Compulsory HBM bytes of k_occlude<K> as (per path with a converged
    item, per converged item).  Per path: the path record (64), K triangles
    and attributes (76 each) and the receiver point (40).  Per item: the list
    entry (4), FP32 solution (8K), wavelength (4) and, written, the throughput
    (4).  BVH nodes and triangles read by the occlusion rays are not counted
    (the BVH stays in cache): the kernel's rate is given by its rays and node
    visits as well."""
    return 64 + 76 * K + 40, 4 + 8 * K + 4 + 4


def FINISH_BYTES(K: int) -> tuple[int, int]:
    """Compulsory HBM bytes of k_finish<K> as (per path with a converged item,
    per converged item).  Per path: path record (64), K triangles and
    attributes (76 each), receiver point (40).  Per item: list entry (4), FP32
    solution (8K), iteration count (4), throughput (4), wavelength (4), the
    48-byte result record and the image word (4)."""
    return 64 + 76 * K + 40, 4 + 8 * K + 4 + 4 + 4 + 48 + 4


def profiled_traffic(workload: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == workload and cells[1] == f"`{kernel}`":
            return float(cells[4]) * 1e6
    return None


def profiled_insts(workload: str, kernel: str):
    """Executed instructions per dispatch of `kernel` from the committed
    rocprofv3 PMC summary (SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_INSTS_SMEM,
    SQ_INSTS_BRANCH: wave instructions), or None."""
    path = os.path.join(ROOT, "profiles", "solver_instruction_mix_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == workload and cells[1] == f"`{kernel}`":
            try:
                return {k: float(v) * 1e9 for k, v in zip(("valu", "salu", "smem", "branch"), cells[2:6])}
            except ValueError:
                return None
    return None


def _solver_flops(stats: dict, event_share: dict) -> dict:
    """Algorithmic FLOPs of the Newton kernels per chain length: residual
    evaluations and steps times the reference iteration's counts, weighted by
    the event mix of that chain length."""
    from palzilla_amd import caustics as CA
    out = {}
    for K in (1, 2, 3):
        row = stats[f"K{K}"]
        share = {ev: w for ev, w in event_share.items() if ev >> 24 == K}
        tot = sum(share.values())
        if not tot:
            out[K] = 0.0
            continue
        f_eval = sum(CA.iteration_flops(ev)[0] * w for ev, w in share.items()) / tot
        f_step = CA.ITERATION_FLOPS_STEP[K]
        steps = row["iterations"] - row["converged"]
        out[K] = row["iterations"] * f_eval + steps * f_step
    return out


def bench_c2(args) -> dict:
    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Event, Stream

    st = Stream()
    t_setup = time.perf_counter()
    s = scenes.config2()
    L = s.lambdas.shape[0]
    log(f"[bench] c2: {s.n_tris} triangles, {s.n_tris * s.photons_per_triangle} photons, "
        f"{s.points.shape[0]} receiver points, {L} wavelengths")
    g = CA.GpuScene(s, stream=st)
    e = [Event() for _ in range(6)]
    e[0].record(st)
    g.trace()
    e[1].record(st)
    info = g.build_table()
    e[2].record(st)
    _, n_all = g.lookup(0, capacity=1)
    d_all, n_all = g.lookup(0, capacity=n_all)
    e[3].record(st)
    st.sync()
    recs_event = g.records()["event"]
    events = sorted(int(v) for v in np.unique(recs_event[recs_event != 0]))
    log(f"[bench] c2: {info['n_tuples']} tuples, {n_all} paths")

    # plain: one list and one launch per event descriptor (pk.h:1480-1586)
    plain_lists = []
    for ev in events:
        K = ev >> 24
        if K > 3:
            continue
        _, n = g.lookup(ev, capacity=1)
        if n:
            d, n = g.lookup(ev, capacity=n)
            plain_lists.append((K, ev, d, n))
    n_plain_paths = sum(x[3] for x in plain_lists)

    bs = CA.BatchedSolver(g.sorter, stream=st)
    e[4].record(st)
    bs.prepare(d_all, n_all)
    e[5].record(st)
    st.sync()
    ranges = bs.ranges()
    assert ranges[3] == n_plain_paths, "batched and plain solve different tuple sets"
    pipeline_ms = {"photon_trace": e[0].elapsed_ms(e[1]), "tuple_table": e[1].elapsed_ms(e[2]),
                   "lookup_flatten": e[2].elapsed_ms(e[3]), "batched_prepare": e[4].elapsed_ms(e[5])}
    n_items = int(ranges[3]) * L
    d_scenes = CA.upload_scene_descs([g.desc()], st)
    d_res = DeviceBuffer(max(int(ranges[5]) * L, 1) * CA.PATH_RESULT.itemsize)

    if args.launch_config:
        cfg = json.loads(args.launch_config)
        for K, c in enumerate(cfg[:3], start=1):
            bs.set_config(K, *c)
        if len(cfg) > 3:
            bs.set_config(0, *cfg[3])
        bs.set_fused(args.fused)
        tune = [bs.get_config(K) for K in (1, 2, 3)]
        log(f"[bench] c2: launch shapes {tune}")
    else:
        log("[bench] c2: autotune")
        tune = bs.autotune(d_scenes, 1, L, refine=CA.REFINE_VERTICES)
        log(f"[bench] c2: tuned {tune}, fused {bs.get_fused()[1]}")

    # batched, timed region: eager launches with per-kernel events on the stream
    for _ in range(args.warmup):
        bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res)
    bs.set_timing(True)
    kern = {k: 0.0 for k in CA.BatchedSolver.TIMED_KERNELS}
    st.sync()
    t0, t1 = Event(), Event()
    log(f"[bench] c2: timing {args.steps} batched steps")
    wall0 = time.perf_counter()
    t0.record(st)
    for _ in range(args.steps):
        bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res)
        for k, v in bs.timings().items():
            kern[k] += v / args.steps
    t1.record(st)
    st.sync()
    wall = time.perf_counter() - wall0
    bs.set_timing(False)
    ms = t0.elapsed_ms(t1) / args.steps
    stats = bs.stats()
    path_stats = bs.path_stats()
    timed_mask = bs.get_fused()[1]  # the structure the timed steps ran

    # the same solve with the per-tuple culling in front of Newton: the
    # interval test (mode 1), then the interval test followed by the affine
    # test (mode 2)
    culling = {"items": n_items}
    for mode, name in ((1, "interval"), (2, "affine")):
        log(f"[bench] c2: one solve with {name} culling")
        bs.set_culling(mode)
        bs.set_timing(True)
        bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res)
        cull_kern = bs.timings()
        cull_stats = bs.stats()
        bs.set_timing(False)
        bs.set_culling(0)
        culling[name] = {"culled_items": cull_stats["culled"], "cull_kernel_ms": round(cull_kern["cull"], 3),
                         "step_ms_with_culling": round(sum(cull_kern.values()), 3),
                         "contributing_items_with_culling":
                             sum(cull_stats[f"K{k}"]["contributing"] for k in (1, 2, 3))}

    # spectral resolve of the receiver images (CIE fit, XYZ -> sRGB, pack):
    # reads L images and the receiver points, writes float4 + RGBA8 per point
    w_l = np.full(L, (float(s.lambdas.max()) - float(s.lambdas.min())) / max(L - 1, 1), dtype=np.float32)
    d_w = DeviceBuffer.from_array(w_l)
    n_pts = s.points.shape[0]
    CA.spectral_resolve(g.d_image, g.d_lambdas, d_w, g.d_points, n_pts, L, stream=st)
    st.sync()
    t0.record(st)
    for _ in range(10):
        CA.spectral_resolve(g.d_image, g.d_lambdas, d_w, g.d_points, n_pts, L, stream=st)
    t1.record(st)
    st.sync()
    ms_resolve = t0.elapsed_ms(t1) / 10
    resolve_bytes = n_pts * (4 * L + 40 + 16 + 4)
    resolve = {"ms": round(ms_resolve, 4), "bytes": resolve_bytes,
               "GBps": round(resolve_bytes / (ms_resolve * 1e-3) / 1e9, 1)}

    # camera side: solvePrimary over a 1024^2 image of the same scene (lens on)
    cam = CA.Camera.lookat((0.2, 3.0, 2.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.9, 1024, 1024)
    ppr = CA.PrimaryParams()
    ppr.p_light = CA.Float3(*s.p_light)
    ppr.light_intensity, ppr.radiance_clamp = 16.0, 12.0
    ppr.cauchy_a, ppr.cauchy_b = s.cauchy_a, s.cauchy_b
    ppr.width, ppr.height, ppr.albedo = 1024, 1024, CA.Float3(0.8, 0.8, 0.8)
    ppr.use_lens, ppr.lens_distance, ppr.lens_thickness, ppr.lens_R = 1, 0.3, 0.05, 0.6
    d_fd, d_acc = CA.solve_primary(g, cam, ppr, stream=st)
    st.sync()
    t0.record(st)
    for it in range(5):
        ppr.iteration = it
        CA.solve_primary(g, cam, ppr, d_fd, d_acc, stream=st)
    t1.record(st)
    st.sync()
    ms_primary = t0.elapsed_ms(t1) / 5
    primary = {"ms": round(ms_primary, 4), "pixels": 1024 * 1024,
               "Mrays_per_s": round(1024 * 1024 / (ms_primary * 1e-3) / 1e6, 1)}

    # the same solve as one hipGraph launch
    g.clear_image()
    bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res, graph=True)
    st.sync()
    t0.record(st)
    for _ in range(args.steps):
        bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res, graph=True)
    t1.record(st)
    st.sync()
    ms_graph = t0.elapsed_ms(t1) / args.steps
    log("[bench] c2: fused mode")
    fused = bench_fused(args, bs, d_scenes, 1, L, d_res, st, graph=False, check=False)
    g.clear_image()
    bs.solve(d_scenes, 1, L, CA.REFINE_VERTICES, d_res)
    img_batched = g.image().reshape(-1).astype(np.float64)

    # plain, same tuple set, image only (as the reference renders)
    log(f"[bench] c2: timing {args.steps} plain steps ({len(plain_lists)} launches each)")
    for _ in range(max(1, args.warmup)):
        for K, ev, d, n in plain_lists:
            g.solve_plain(K, ev, d, n)
    pe = [(Event(), Event()) for _ in plain_lists]
    plain_kernel_ms = {}
    g.clear_image()
    st.sync()
    t0.record(st)
    for _ in range(args.steps):
        for (K, ev, d, n), (a, b) in zip(plain_lists, pe):
            a.record(st)
            g.solve_plain(K, ev, d, n)
            b.record(st)
        st.sync()
        for (K, ev, d, n), (a, b) in zip(plain_lists, pe):
            plain_kernel_ms[CA.event_str(ev)] = plain_kernel_ms.get(CA.event_str(ev), 0.0) + \
                a.elapsed_ms(b) / args.steps
    t1.record(st)
    st.sync()
    ms_plain = t0.elapsed_ms(t1) / args.steps
    g.clear_image()
    for K, ev, d, n in plain_lists:
        g.solve_plain(K, ev, d, n)
    img_plain = g.image().reshape(-1).astype(np.float64)
    st.sync()
    img_rmse = float(np.sqrt(np.mean((img_batched - img_plain) ** 2)) /
                     max(np.sqrt(np.mean(img_plain ** 2)), 1e-30))

    converged = sum(stats[f"K{k}"]["converged"] for k in (1, 2, 3))
    contributing = sum(stats[f"K{k}"]["contributing"] for k in (1, 2, 3))

    # sample: CPU oracle (baseline + FP32 parity) and FP64 vertex error of the refined chains
    sample = verify_sample(args, s, g, bs, d_res, n_all, L, st)
    frac_1e5 = sample["vertex"]["frac_within_1e-5"]
    value = contributing * frac_1e5 / (ms * 1e-3)
    # the plain kernel's chains are the FP32 ones (bit-identical to the oracle's)
    plain_value = contributing * sample["vertex"]["fp32_frac_within_1e-5"] / (ms_plain * 1e-3)

    # roofline of the dominant kernel.  The path uses no MFMA; the roofline is
    # the HBM one over each kernel's compulsory bytes per item (FINISH_BYTES /
    # NEWTON_BYTES), next to the Newton kernels' FP32 rate over the reference
    # iteration's FLOPs.
    event_share = {}
    for K, ev, d, n in plain_lists:
        event_share[ev] = event_share.get(ev, 0) + n
    flops = _solver_flops(stats, event_share)
    # the kernels of the solve as the autotuner structured it (fused mask bit 1:
    # one Newton kernel over every chain length, bit 2: one occlusion kernel),
    # each with its compulsory bytes: (paths, bytes per path, items, bytes per
    # item) per chain length it covers
    mask = timed_mask
    waves = bs.get_config(0)["chunk"]
    kernels = {}  # key -> (kernel name, ms, [(K, paths, bytes/path, items, bytes/item)])

    def newton_units(K):
        return (K, int(ranges[K]) - int(ranges[K - 1]), NEWTON_BYTES(K)[0], stats[f"K{K}"]["items"],
                NEWTON_BYTES(K)[1])

    def occlude_units(K):
        return (K, path_stats[f"K{K}"], OCCLUDE_BYTES(K)[0], stats[f"K{K}"]["converged"], OCCLUDE_BYTES(K)[1])
    if mask & 1:
        kernels["newton_mixed"] = (f"k_newton_mixed<{waves}>", kern["newton_k1"], [newton_units(K) for K in (1, 2, 3)])
    else:
        for K in (1, 2, 3):
            kernels[f"newton_k{K}"] = (f"k_newton<{K}>", kern[f"newton_k{K}"], [newton_units(K)])
    if mask & 2:
        kernels["occlude_mixed"] = ("k_occlude_mixed", kern["occlude_k1"], [occlude_units(K) for K in (1, 2, 3)])
    else:
        for K in (1, 2, 3):
            kernels[f"occlude_k{K}"] = (f"k_occlude<{K}>", kern[f"occlude_k{K}"], [occlude_units(K)])
    for K in (1, 2, 3):
        kernels[f"finish_k{K}"] = (f"k_finish<{K}, 1>", kern[f"finish_k{K}"],
                                   [(K, path_stats[f"K{K}"], FINISH_BYTES(K)[0], stats[f"K{K}"]["converged"],
                                     FINISH_BYTES(K)[1])])
    kernels["compact"] = ("k_compact_scatter", kern["compact"], None)
    dominant = max(kernels, key=lambda k: kernels[k][1])
    name, dom_ms, dom_units = kernels[dominant]
    roof = {"bound": "hbm", "kernel": name, "kernel_ms": round(dom_ms, 3), "fused_mask": mask}
    if dom_units:
        algo = sum(n_p * per_p + n_i * per_i for _, n_p, per_p, n_i, per_i in dom_units)
        achieved = algo / (dom_ms * 1e-3) / 1e9
        roof.update({"achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "algorithmic_bytes": algo,
                     "units": {f"K{K}": {"paths": n_p, "bytes_per_path": per_p, "items": n_i, "bytes_per_item": per_i}
                               for K, n_p, per_p, n_i, per_i in dom_units}})
    roof["traffic"] = profiled_traffic("c2", name)
    if dominant.startswith("newton"):
        # The Newton kernels are compute-bound (their HBM rate is the `hbm`
        # entry below): the roof is the FP32 peak over the reference
        # iteration's FLOPs.  The path has no matrix-shaped work, so the FP32
        # VALU pipe is the one that runs; its executed instructions per
        # dispatch come from the committed PMC summary.
        f = sum(flops[K] for K, *_ in dom_units)
        tflops = f / (dom_ms * 1e-3) / 1e12
        roof["hbm"] = {k: roof.pop(k) for k in ("achieved", "peak", "unit", "frac", "algorithmic_bytes") if k in roof}
        roof.update({"bound": "mfma", "pipe": "FP32 VALU", "achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "algorithmic_flops": f,
                     "counted": "reference-equivalent FLOPs (2K single-tangent passes per iteration)"})
        insts = profiled_insts("c2", name)
        if insts:
            roof["executed"] = {"valu_wave_insts": insts["valu"], "salu_wave_insts": insts["salu"],
                                "valu_G_per_s": round(insts["valu"] / (dom_ms * 1e-3) / 1e9, 1)}
    # the occlusion kernels are traversal-bound: rays per second and the
    # escape-link walk's work per ray (oracle counts on the sample)
    ow = sample.get("occlusion_work")
    if ow and ow["converged_items"]:
        rays_per_item = ow["rays"] / ow["converged_items"]
        traversal = {}
        for key, (kname, t_ms, kunits) in kernels.items():
            if not key.startswith("occlude"):
                continue
            conv = sum(n_i for _, _, _, n_i, _ in kunits)
            if t_ms > 0 and conv:
                traversal[kname] = {"converged_items_per_s": round(conv / (t_ms * 1e-3), 1),
                                    "rays_per_s_est": round(conv * rays_per_item / (t_ms * 1e-3), 1)}
        traversal["sample"] = {"rays_per_converged_item": round(rays_per_item, 3),
                               "boxes_per_ray": round(ow["boxes"] / max(ow["rays"], 1), 2),
                               "triangles_per_ray": round(ow["triangles"] / max(ow["rays"], 1), 2),
                               "walk": "escape-link binary walk (minimum_lbvh.h:1078-1152), oracle counts",
                               "paths": ow["paths"]}
        roof["traversal"] = traversal
    valu = {}
    for key, (kname, t_ms, kunits) in kernels.items():
        if not key.startswith("newton"):
            continue
        f = sum(flops[K] for K, *_ in kunits)
        if t_ms > 0 and f > 0:
            valu[kname] = {"achieved_TFLOPs": round(f / (t_ms * 1e-3) / 1e12, 2), "peak_TFLOPs": FP32_PEAK_TFLOPS,
                           "frac": round(f / (t_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4), "flops": f,
                           "counted": "reference-equivalent FLOPs (2K single-tangent passes per iteration)"}
    trace = sample.get("photon_trace_cpu")
    if trace:
        # photonTrace's algorithmic bytes (64 per node visited, 36 per triangle
        # tested, 40 per attribute record read, 56 per tuple stored), from the
        # sampled triangles' counts scaled by the sampling stride
        w = trace["work"]
        per_photon = (64 * w["nodes"] + 36 * w["triangles"] + 40 * w["attributes"] + 56 * w["tuples"]) / max(
            w["photons"], 1)
        total = per_photon * w["photons"] * trace["stride"]
        gbps = total / (pipeline_ms["photon_trace"] * 1e-3) / 1e9
        trace["gpu_roofline"] = {"bound": "hbm (BVH of 8.4 MB largely cache-resident)", "achieved": round(gbps, 1),
                                 "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
                                 "bytes_per_photon": round(per_photon, 1), "photons_traced": w["photons"] * trace["stride"],
                                 "kernel_ms": round(pipeline_ms["photon_trace"], 3)}
    setup_s = time.perf_counter() - t_setup
    return {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "paths/s",
        "ms_per_step": round(ms, 3),
        "dtype": "f32+f64",
        "config": {"workload": "spectral slab: 256x256-cell wavy glass top + flat bottom (131,076 tris), "
                               "2^22 photons, 1024^2 receiver, 16 wavelengths 380-720 nm, TT + all other "
                               "chains of length 1..3; batched FP32 Newton + one FP64 refinement step",
                   "n_triangles": s.n_tris, "photons": s.n_tris * s.photons_per_triangle,
                   "receiver": [s.width, s.height], "wavelengths": L, "paths": int(ranges[3]),
                   "items": n_items, "refine": "fp64 step on converged vertices",
                   "counted": "contributing chains (cost < 1e-10, throughput > 0) x fraction of the "
                              "sampled chains within 1e-5 relative vertex error of the FP64 oracle"},
        "roofline": roof,
        "cpu_baseline": sample["cpu_baseline"],
        "batched": {"ms_per_step": round(ms, 3), "ms_per_step_graph": round(ms_graph, 3),
                    "kernel_ms": {k: round(v, 3) for k, v in kern.items()},
                    "converged_items": converged, "contributing_items": contributing,
                    "stats": stats, "launch_configs": tune, "wall_s": round(wall, 3),
                    "newton_fp32_roofline": valu, "fused": fused,
                    "fused_mask": timed_mask},
        "plain": {"value": round(plain_value, 1), "unit": "paths/s", "ms_per_step": round(ms_plain, 3),
                  "launches": len(plain_lists), "kernel_ms": {k: round(v, 3) for k, v in plain_kernel_ms.items()},
                  "batched_over_plain": round(value / plain_value, 2) if plain_value else None,
                  "batched_over_plain_time": round(ms_plain / ms, 2),
                  "image_rel_rmse_batched_vs_plain": img_rmse},
        "culling": culling,
        "spectral_resolve": resolve,
        "solve_primary": primary,
        "vertex_error": sample["vertex"],
        "parity_sample": sample["parity"],
        "pipeline_ms": {k: round(v, 3) for k, v in pipeline_ms.items()},
        "photon_trace": sample.get("photon_trace_cpu"),
        "setup_s": round(setup_s, 1),
    }


def verify_sample(args, s, g, bs, d_res, n_all, L, st) -> dict:
    """A bounded, evenly spread sample of the path list: the CPU oracle solves
    it (the CPU baseline; its FP32 results must equal the GPU's), and the FP64
    oracle measures the vertex error of the GPU's refined chains."""
    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    import oracle

    n_chunks = 64
    chunk = max(1, min(args.cpu_paths // n_chunks, n_all // n_chunks))
    starts = np.linspace(0, n_all - chunk, n_chunks).astype(np.int64)
    _, p_sorted, _ = bs.pointers()
    paths, res = [], []
    for a in starts:
        a = int(a)
        paths.append(CA.download_paths_at(p_sorted, a, chunk))
        res.append(d_res.download(CA.PATH_RESULT, chunk * L, offset_bytes=a * L * CA.PATH_RESULT.itemsize))
    paths = np.concatenate(paths)
    res = np.concatenate(res)
    keep = (paths["event"] >> 24) <= 3
    paths, res = paths[keep], res[np.repeat(keep, L)]
    log(f"[bench] cpu baseline: oracle solve of {paths.shape[0]} paths x {L} wavelengths, {CPU_THREADS} threads")
    nodes = g.bvh.download_nodes()
    root = g.bvh.download_root()
    hs = oracle.HostScene(s, nodes, root)
    t0 = time.perf_counter()
    want = oracle.solve_paths(hs, paths, threads=CPU_THREADS)
    dt = time.perf_counter() - t0
    cpu_contrib = int((want["status"] == CA.PATH_CONTRIBUTES).sum())
    # the same solve on one thread, over every 8th path of the sample
    sub = paths[::8]
    hs1 = oracle.HostScene(s, nodes, root)
    t0 = time.perf_counter()
    want1 = oracle.solve_paths(hs1, sub, threads=1)
    dt1 = time.perf_counter() - t0
    one_thread = int((want1["status"] == CA.PATH_CONTRIBUTES).sum()) / dt1
    # FP32 parity of the sample: decisions and irradiance equal; params are refined on the GPU
    same = {f: bool(np.array_equal(res[f], want[f])) for f in ("status", "iterations", "contribution",
                                                                 "footprint")}
    conv = np.nonzero(want["status"] >= 1)[0]
    rng = np.random.default_rng(7)
    pick = conv if conv.size <= args.vertex_sample else rng.choice(conv, args.vertex_sample, replace=False)
    e32, e64 = [], []
    vscale = scenes.vertex_error_scale(s)  # the specular geometry's diagonal
    for it in pick:
        pi, l = divmod(int(it), L)
        path = paths[pi]
        K = int(path["event"]) >> 24
        lam = np.float32(s.lambdas[l])
        eta = float(np.float32(s.cauchy_a) + np.float32(s.cauchy_b) / (lam * lam))
        p_end = s.points[path["pixel"]]["p"]
        ok, _, x64 = oracle.solve_f64(s, path, p_end, eta, want["params"][it][:2 * K].astype(np.float64),
                                      tol=1e-28, max_iter=64)
        v = s.tris[path["tris"][:K]].astype(np.float64).reshape(K, 3, 3)

        def verts(prm):
            p = np.asarray(prm, dtype=np.float64)[:2 * K].reshape(K, 2)
            return v[:, 0] + (v[:, 1] - v[:, 0]) * p[:, :1] + (v[:, 2] - v[:, 0]) * p[:, 1:]
        sc = vscale
        e32.append(np.max(np.linalg.norm(verts(want["params"][it]) - x64, axis=1) / sc))
        e64.append(np.max(np.linalg.norm(verts(res["params"][it]) - x64, axis=1) / sc))
    e32, e64 = np.asarray(e32), np.asarray(e64)
    # occlusion-ray work of the solve (oracle counts, serial) on every 64th path of the sample
    osub = paths[::64]
    ow = oracle.occlusion_work(oracle.HostScene(s, nodes, root), osub)
    ow["paths"] = int(osub.shape[0])
    ow["converged_items"] = int((want["status"].reshape(-1, L)[::64] >= 1).sum())
    # photonTrace on the CPU (the oracle's stackfull traversal, 1 thread) over
    # every 64th triangle's photons, with its node and triangle counts
    stride = 64
    t0 = time.perf_counter()
    work = oracle.photon_trace_work(s, nodes, root, stride=stride)
    dt_trace = time.perf_counter() - t0
    return {
        "occlusion_work": ow,
        "photon_trace_cpu": {"value": round(work["photons"] / dt_trace, 1), "unit": "photons/s", "cores": 1,
                             "kind": "port", "sample": f"oracle photonTrace of every {stride}th triangle's photons "
                                                       f"({work['photons']} photons, {dt_trace:.2f} s)",
                             "work": work, "stride": stride},
        "cpu_baseline": {"value": round(cpu_contrib / dt, 1), "unit": "paths/s", "cores": CPU_THREADS,
                         "kind": "port", "host_cores": os.cpu_count(), "value_1_thread": round(one_thread, 1),
                         "sample_1_thread": f"every 8th path of the sample, {sub.shape[0]} paths, {dt1:.2f} s",
                         "sample": f"oracle FP32 solve (solveConstraints + throughput + dA/dw) of "
                                   f"{paths.shape[0]} paths x {L} wavelengths in {n_chunks} evenly spaced "
                                   f"chunks of the path list, {CPU_THREADS} threads, {dt:.1f} s"},
        "parity": {"items": int(res.shape[0]), "equal_fields": same,
                   "contributing": cpu_contrib},
        "vertex": {"chains": int(pick.size), "fp32_max": float(e32.max()) if e32.size else None,
                   "fp32_p99": float(np.quantile(e32, 0.99)) if e32.size else None,
                   "refined_max": float(e64.max()) if e64.size else None,
                   "refined_p99": float(np.quantile(e64, 0.99)) if e64.size else None,
                   "frac_within_1e-5": float(np.mean(e64 <= 1e-5)) if e64.size else 0.0,
                   "fp32_frac_within_1e-5": float(np.mean(e32 <= 1e-5)) if e32.size else 0.0},
    }


# ---------------------------------------------------------------------------------------------------
# 256 scenes, batched vs per-scene plain

def bench_c4(args) -> dict:
    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Event, Stream

    st = Stream()
    scene_list = scenes.config4_scenes(n_scenes=args.c4_scenes)
    log(f"[bench] c4: {len(scene_list)} scenes, {sum(x.n_tris for x in scene_list)} triangles")
    batch = CA.SceneBatch(scene_list, stream=st)
    n = batch.build()
    L = batch.n_lambdas
    st.sync()
    # per-scene, per-descriptor lists for the plain launches
    plain = []
    for i, gs in enumerate(batch.scenes):
        ev_all = gs.records()["event"]
        for ev in sorted(int(v) for v in np.unique(ev_all[ev_all != 0])):
            if ev >> 24 > 3:
                continue
            _, m = gs.lookup(ev, capacity=1, scene_index=i)
            if m:
                d, m = gs.lookup(ev, capacity=m, scene_index=i)
                plain.append((gs, ev >> 24, ev, d, m))
    bs = CA.BatchedSolver(batch.sorter, stream=st)
    bs.prepare(batch.d_paths, n)
    ranges = bs.ranges()
    assert ranges[3] == sum(x[4] for x in plain)
    if args.launch_config:
        # fixed shapes (a profiling run): per length and, when given, the fused kernels'
        cfg = json.loads(args.launch_config)
        for K, c in enumerate(cfg[:3], start=1):
            bs.set_config(K, *c)
        if len(cfg) > 3:
            bs.set_config(0, *cfg[3])
        bs.set_fused(args.fused if args.fused >= 0 else (3 if len(cfg) > 3 else 0))
        tune = [bs.get_config(K) for K in (1, 2, 3)]
    else:
        tune = bs.autotune(batch.d_scenes, len(scene_list), L, refine=CA.REFINE_VERTICES)
    fused_chosen = bs.get_fused()[1]
    d_res = DeviceBuffer(max(n * L, 1) * CA.PATH_RESULT.itemsize)
    t0, t1 = Event(), Event()
    for _ in range(args.warmup + 1):
        bs.solve(batch.d_scenes, len(scene_list), L, CA.REFINE_VERTICES, d_res, graph=True)
    st.sync()
    t0.record(st)
    for _ in range(args.steps):
        bs.solve(batch.d_scenes, len(scene_list), L, CA.REFINE_VERTICES, d_res, graph=True)
    t1.record(st)
    st.sync()
    ms = t0.elapsed_ms(t1) / args.steps
    stats = bs.stats()
    contributing = sum(stats[f"K{k}"]["contributing"] for k in (1, 2, 3))
    fused = bench_fused(args, bs, batch.d_scenes, len(scene_list), L, d_res, st, graph=True, check=True)
    for _ in range(max(1, args.warmup)):
        for gs, K, ev, d, m in plain:
            gs.solve_plain(K, ev, d, m)
    st.sync()
    t0.record(st)
    for _ in range(args.steps):
        for gs, K, ev, d, m in plain:
            gs.solve_plain(K, ev, d, m)
    t1.record(st)
    st.sync()
    ms_plain = t0.elapsed_ms(t1) / args.steps
    # parity at this size: the batched solve and the plain launches sum the same
    # contributions into per-scene fixed-point images, which must be equal bit for bit
    fixed_b = [DeviceBuffer(max(g.n_points * L, 1) * 8) for g in batch.scenes]
    fixed_p = [DeviceBuffer(max(g.n_points * L, 1) * 8) for g in batch.scenes]
    for b in fixed_b + fixed_p:
        b.zero(st)
    d_fx = CA.upload_scene_descs([g.desc(image_fixed=b.ptr) for g, b in zip(batch.scenes, fixed_b)], st)
    bs.solve(d_fx, len(scene_list), L, CA.REFINE_VERTICES, d_res)
    index = {id(g): i for i, g in enumerate(batch.scenes)}
    for gs, K, ev, d, m in plain:
        gs.solve_plain(K, ev, d, m, image_fixed=fixed_p[index[id(gs)]].ptr)
    st.sync()
    equal = all(a.download(np.uint8, a.nbytes, st).tobytes() == b.download(np.uint8, b.nbytes, st).tobytes()
                for a, b in zip(fixed_b, fixed_p))
    return {"workload": "256 stack scenes, 16^2..256^2 cells, K = 1..3, IOR U[1.3,1.8], 2^14..2^20 photons",
            "images_equal_batched_vs_plain": equal,
            "paths": int(ranges[3]), "items": int(ranges[3]) * L, "contributing_items": contributing,
            "batched_graph_ms": round(ms, 3), "plain_ms": round(ms_plain, 3), "plain_launches": len(plain),
            "batched_paths_per_s": round(contributing / (ms * 1e-3), 1),
            "plain_paths_per_s": round(contributing / (ms_plain * 1e-3), 1),
            "batched_over_plain": round(ms_plain / ms, 2), "launch_configs": tune,
            "fused_chosen_by_autotuner": fused_chosen, "fused": fused}


def bench_fused(args, bs, d_scenes, n_scenes, L, d_res, st, graph: bool, check: bool) -> dict:
    """The same solve in fused mode (one persistent Newton kernel and one
    occlusion kernel over every chain length, shapes from the autotuner):
    time per step, per-kernel times of eager solves and, with `check`, whether
    every per-item result equals the per-length kernels' bytes (d_res holds
    those; without `check` it is overwritten)."""
    from palzilla_amd import caustics as CA
    from palzilla_amd.device import DeviceBuffer, Event
    ref = CA.download_results(d_res, int(bs.ranges()[5]) * L) if check else None
    prev_mode = bs.get_fused()[0]
    bs.set_fused(True)
    try:
        cfg = json.loads(args.launch_config) if args.launch_config else []
        if len(cfg) > 3:
            bs.set_config(0, *cfg[3])
        else:
            bs.autotune(d_scenes, n_scenes, L, refine=CA.REFINE_VERTICES)
        shape = bs.get_config(0)
        d_out = DeviceBuffer(d_res.nbytes) if check else d_res
        for _ in range(args.warmup + 1):
            bs.solve(d_scenes, n_scenes, L, CA.REFINE_VERTICES, d_out, graph=graph)
        t0, t1 = Event(), Event()
        st.sync()
        t0.record(st)
        for _ in range(args.steps):
            bs.solve(d_scenes, n_scenes, L, CA.REFINE_VERTICES, d_out, graph=graph)
        t1.record(st)
        st.sync()
        ms = t0.elapsed_ms(t1) / args.steps
        bs.set_timing(True)
        bs.solve(d_scenes, n_scenes, L, CA.REFINE_VERTICES, d_out)
        k = bs.timings()
        bs.set_timing(False)
        same = None
        if ref is not None:
            same = CA.download_results(d_out, ref.shape[0]).tobytes() == ref.tobytes()
    finally:
        bs.set_fused(prev_mode)
    return {"ms_per_step": round(ms, 3), "graph": graph, "results_equal_per_length_kernels": same,
            "newton_mixed_ms": round(k["newton_k1"], 3), "occlude_mixed_ms": round(k["occlude_k1"], 3),
            "finish_ms": {f"K{K}": round(k[f"finish_k{K}"], 3) for K in (1, 2, 3)},
            "compact_ms": round(k["compact"], 3), "shape": shape}


# ---------------------------------------------------------------------------------------------------
# adjoint: d(L2 image loss)/d(heights) of a 128^2-cell refractive heightfield

def bench_path_cuts(args) -> dict:
    """Node-level path cuts (palz_path_cuts_enumerate) of TT chains on a 16^2-cell
    slab (516 triangles) to a 32^2 receiver, 16 wavelengths: count-only and
    full (count + records) calls, timed with the stream synchronised."""
    import time

    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Stream

    s = scenes.config2(cells=16, width=32, photons=1 << 14, n_lambdas=16)
    st = Stream()
    g = CA.GpuScene(s, stream=st)
    d_nb = g.bvh.normal_bounds(g.d_attribs.ptr, st)
    pc = CA.PathCuts(stream=st)
    event, eta = CA.event_u32("TT"), CA.eta_range(s)
    desc = g.desc()
    _, n = pc.enumerate(desc, d_nb, event, *eta, capacity=0)
    st.sync()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        _, n = pc.enumerate(desc, d_nb, event, *eta, capacity=0)
    st.sync()
    ms_count = (time.perf_counter() - t) * 1e3 / reps
    d_paths = DeviceBuffer(max(n, 1) * CA.PATH.itemsize)
    pc.enumerate(desc, d_nb, event, *eta, capacity=n, d_paths=d_paths)
    st.sync()
    t = time.perf_counter()
    for _ in range(reps):
        pc.enumerate(desc, d_nb, event, *eta, capacity=n, d_paths=d_paths)
    st.sync()
    ms_full = (time.perf_counter() - t) * 1e3 / reps
    assert pc.status() == 0
    out = {"workload": "TT path cuts, 16^2-cell slab (516 tris), 32^2 receiver points, eta over 16 wavelengths",
           "tuples": n, "count_only_ms": round(ms_count, 2), "full_ms": round(ms_full, 2),
           "tuples_per_s": round(n / (ms_full * 1e-3), 1), "record_bytes": n * CA.PATH.itemsize}
    if not args.no_cpu:
        import oracle
        nodes, root = g.bvh.download_nodes(), g.bvh.download_root()
        nb = oracle.normal_bounds(nodes, s.attribs)
        pts = s.points[[0, 341, 682, 1023]]
        t = time.perf_counter()
        _, tup = oracle.path_cuts(s.tris, s.attribs, nodes, root, nb, pts, s.p_light, event, *eta)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(tup.shape[0] / dt, 1), "unit": "tuples/s", "cores": 1, "kind": "port",
                               "sample": f"oracle traversal of 4 receiver points ({tup.shape[0]} tuples, {dt:.1f} s)"}
    return out


def bench_renderer(args) -> dict:
    """The library's PKRenderer-shaped driver (palz_renderer_step) on the C2
    geometry seen by a 512^2 camera (lens on), 16 wavelengths, 32 photons per
    triangle (4.2M photons): whole frames, stream-synchronised."""
    import time

    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    from palzilla_amd.device import Stream

    s = scenes.config2(cells=256, width=8, photons=1 << 22, n_lambdas=16)
    st = Stream()
    lam = s.lambdas.astype(np.float32)
    w = np.full(lam.shape[0], (float(lam.max()) - float(lam.min())) / (lam.shape[0] - 1), dtype=np.float32)
    W = H = 512
    r = CA.Renderer(W, H, lam, w, stream=st)
    p = r.params()
    p.p_light = CA.Float3(*s.p_light)
    p.cauchy_a, p.cauchy_b = s.cauchy_a, s.cauchy_b
    p.photons_per_triangle = s.photons_per_triangle
    p.use_lens, p.lens_distance, p.lens_thickness, p.lens_R = 1, 0.3, 0.05, 0.6
    p.albedo = CA.Float3(0.8, 0.8, 0.8)
    p.path_capacity = 1 << 26
    r.set_params(p)
    r.set_camera(CA.Camera.lookat((0.2, 3.0, 2.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.9, W, H))
    r.load_scene(s.tris, s.attribs)
    r.step()
    st.sync()
    frames = 3
    t = time.perf_counter()
    for _ in range(frames):
        r.step()
    st.sync()
    ms = (time.perf_counter() - t) * 1e3 / frames
    _, _, counts = r.buffers()
    pix = r.resolve()
    return {"workload": "C2 slab (131,076 tris) seen by a 512^2 camera with the thick lens, 16 wavelengths, "
                        "32 photons per triangle (4.2M), every chain of length 1-3 batched + T,T,T,T plain",
            "ms_per_frame": round(ms, 2), "frames_per_s": round(1e3 / ms, 2), "paths_last_frame": int(counts[1]),
            "lit_pixels": int(np.count_nonzero(pix & 0x00FFFFFF))}


def bench_c5(args) -> dict:
    from palzilla_amd import caustics as CA
    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Event, Stream

    cells, width = 128, 512
    st = Stream()
    h = scenes.slab_heights(cells, 5)

    def forward(heights, g_prev=None):
        s = scenes.config5(heights, cells=cells, width=width)
        g = CA.GpuScene(s, stream=st)
        return s, g

    # target image from a second seed
    s_t, g_t = forward(scenes.slab_heights(cells, 6))
    g_t.trace()
    g_t.build_table()
    _, n_t = g_t.lookup(0, capacity=1)
    d_pt, n_t = g_t.lookup(0, capacity=n_t)
    bt = CA.BatchedSolver(g_t.sorter, stream=st)
    bt.prepare(d_pt, n_t)
    g_t.clear_image()
    bt.solve(CA.upload_scene_descs([g_t.desc()], st), 1, 1, CA.REFINE_VERTICES, None)
    target = g_t.image().reshape(-1)

    s, g = forward(h)
    tv, nb, coef = scenes.height_param_tables(cells, s.n_tris)
    hp = CA.HeightParams(tv, nb, coef)
    e = [Event() for _ in range(4)]
    ms_fwd, ms_bwd = [], []
    # sizes of the path list, so that the timed steps allocate nothing
    g.trace()
    g.build_table()
    _, cap = g.lookup(0, capacity=1)
    cap = int(cap * 1.1) + 1024
    d_paths = DeviceBuffer(cap * CA.PATH.itemsize)
    bs = CA.BatchedSolver(g.sorter, stream=st)
    d_scenes = CA.upload_scene_descs([g.desc()], st)
    d_res = DeviceBuffer(cap * CA.PATH_RESULT.itemsize)
    d_g = DeviceBuffer(hp.n_params * 8)
    for step in range(args.warmup + args.steps):
        e[0].record(st)
        g.trace()
        g.build_table()
        d_paths, n = g.lookup(0, capacity=cap, d_paths=d_paths)
        bs.prepare(d_paths, min(n, cap))
        g.clear_image()
        bs.solve(d_scenes, 1, 1, CA.REFINE_VERTICES, d_res)
        e[1].record(st)
        img = g.image().reshape(-1)
        weight = (2.0 * (img - target)).astype(np.float32)
        d_w = DeviceBuffer.from_array(weight, st)
        d_g.zero(st)
        _, p_sorted, _ = bs.pointers()
        e[2].record(st)
        CA.adjoint_irradiance(g, hp, p_sorted, min(n, cap), d_res.ptr, d_w, d_g, st)
        e[3].record(st)
        st.sync()
        if step >= args.warmup:
            ms_fwd.append(e[0].elapsed_ms(e[1]))
            ms_bwd.append(e[2].elapsed_ms(e[3]))
    grad = d_g.download(np.float64, hp.n_params, st)
    loss = float(np.sum((img.astype(np.float64) - target) ** 2))
    fwd, bwd = float(np.median(ms_fwd)), float(np.median(ms_bwd))
    return {"workload": "adjoint: 128^2-cell refractive heightfield (16,641 vertex heights), IOR 1.5, "
                        "512^2 receiver, L2 loss to a second seed's image",
            "paths": int(n), "params": int(hp.n_params), "loss": loss, "grad_norm": float(np.linalg.norm(grad)),
            "forward_ms": round(fwd, 3), "backward_ms": round(bwd, 3), "backward_over_forward": round(bwd / fwd, 3),
            "forward_includes": "photon trace, tuple table, flatten, batched solve (FP32 + FP64 step), image"}


# ---------------------------------------------------------------------------------------------------
# LBVH + onesweep

def bench_lbvh(args) -> dict:
    from palzilla_amd import scenes
    from palzilla_amd.device import DeviceBuffer, Event, Stream, sync
    from palzilla_amd.lbvh import BVHGPUBuilder, OnesweepSort

    n = args.n
    log(f"[bench] lbvh: generating {n} triangles")
    tris = scenes.triangle_soup(n, seed=2024)
    d_tris = DeviceBuffer.from_array(tris)
    st = Stream()
    sorter = OnesweepSort()
    sorter.reserve(n)
    b = BVHGPUBuilder()
    for _ in range(args.warmup):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
    st.sync()
    sync()
    e0, e1 = Event(), Event()
    e0.record(st)
    for _ in range(args.steps):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
    e1.record(st)
    st.sync()
    ms = e0.elapsed_ms(e1) / args.steps
    b.set_timing(True)
    stage_ms = {k: 0.0 for k in b.STAGES}
    for _ in range(args.steps):
        b.build(d_tris.ptr, n, 0, sorter, stream=st, morton_bits=30)
        for k, v in b.timings().items():
            stage_ms[k] += v / args.steps
    b.set_timing(False)
    words, _ = b.download_sorted()
    keys = (words & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.all(np.diff(keys) >= 0), "sorted keys are not sorted"
    assert np.array_equal(np.sort(words >> np.uint64(32)), np.arange(n, dtype=np.uint64))
    # the smaller sizes of the same workload (1M and 4M triangles): time per build
    by_size = {}
    for m in (1 << 20, 1 << 22):
        if m >= n:
            continue
        for _ in range(args.warmup):
            b.build(d_tris.ptr, m, 0, sorter, stream=st, morton_bits=30)
        e0.record(st)
        for _ in range(args.steps):
            b.build(d_tris.ptr, m, 0, sorter, stream=st, morton_bits=30)
        e1.record(st)
        st.sync()
        ms_m = e0.elapsed_ms(e1) / args.steps
        by_size[str(m)] = {"ms_per_build": round(ms_m, 4), "Mtri_per_s": round(m / (ms_m * 1e-3) / 1e6, 1)}
    achieved = LBVH_BYTES_PER_TRI * n / (ms * 1e-3) / 1e9
    copy_gbps = measure_copy_bandwidth()
    ref_bytes = LBVH_REF_PASSES_BYTES[0] * n + LBVH_REF_PASSES_BYTES[1] * (n - 1)
    ref_gbps = ref_bytes / (ms * 1e-3) / 1e9
    out = {"workload": "lbvh+onesweep, 16M-triangle soup (30-bit Morton keys)", "n_triangles": n,
           "value": round(n / (ms * 1e-3) / 1e6, 2), "unit": "Mtri/s", "ms_per_step": round(ms, 4),
           "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
           "algorithmic_bytes_per_tri": LBVH_BYTES_PER_TRI, "achieved_GBps": round(achieved, 1),
           "hbm_frac": round(achieved / HBM_PEAK_GBPS, 4),
           "stage_GBps": {k: round(STAGE_BYTES_PER_TRI[k] * n / (v * 1e-3) / 1e9, 1)
                          for k, v in stage_ms.items() if v > 0 and STAGE_BYTES_PER_TRI[k] > 0},
           "copy_kernel_GBps": round(copy_gbps, 1),
           "reference_pass_bytes": {"bytes": ref_bytes, "GBps": round(ref_gbps, 1),
                                    "frac_of_8TBps": round(ref_gbps / HBM_PEAK_GBPS, 4),
                                    "frac_of_copy_kernel": round(ref_gbps / copy_gbps, 4)},
           "smaller_sizes": by_size}
    if not args.no_cpu:
        import oracle
        sample = min(n, args.cpu_sample)
        t0 = time.perf_counter()
        oracle.lbvh(tris[:sample], bits=30)
        dt1 = time.perf_counter() - t0
        # all host cores: the BVHCPUBuilder shape with CPU_PARALLEL (threaded
        # radix sort, atomic climb), over the whole build
        big = min(n, 4 * args.cpu_sample)
        t0 = time.perf_counter()
        oracle.lbvh_parallel(tris[:big], bits=30, threads=CPU_THREADS)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(big / dt / 1e6, 3), "unit": "Mtri/s", "cores": CPU_THREADS,
                               "kind": "port", "host_cores": os.cpu_count(),
                               "sample": f"oracle LBVH build on {CPU_THREADS} threads (scene AABB, Morton keys, "
                                         f"stable radix sort, deltas, atomic bottom-up climb) of the first {big} "
                                         f"triangles, {dt:.2f} s",
                               "value_1_thread": round(sample / dt1 / 1e6, 3),
                               "sample_1_thread": f"serial oracle build (std::stable_sort, serial climb) of the "
                                                  f"first {sample} triangles, {dt1:.2f} s"}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=["c2", "c4", "c5", "lbvh", "pathcuts", "renderer"])
    ap.add_argument("--no-extras", action="store_true", help="c2 only, without the c4 and lbvh sub-results")
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--cpu-sample", type=int, default=1 << 22, help="LBVH CPU baseline triangles")
    ap.add_argument("--cpu-paths", type=int, default=1 << 21, help="c2 CPU baseline paths")
    ap.add_argument("--vertex-sample", type=int, default=512)
    ap.add_argument("--c4-scenes", type=int, default=256)
    ap.add_argument("--launch-config", default=None,
                    help="skip the c2 autotuner: JSON [[block, blocks_per_cu, chunk, finish_block]] for K = 1..3 "
                         "and, optionally, the fused kernels' [block, blocks_per_cu, waves per SIMD, occlusion block]")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fused", type=int, default=-1,
                    help="with --launch-config: the fused mask (1 Newton, 2 occlusion, 3 both; -1: per length "
                         "for c2, as the shapes say for c4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); nothing is run across GPUs.",
              file=sys.stderr)
        return 2
    if args.workload == "c2":
        out = bench_c2(args)
        if not args.no_extras:
            out["c4"] = bench_c4(args)
            out["c5"] = bench_c5(args)
            out["lbvh"] = bench_lbvh(args)
            out["path_cuts"] = bench_path_cuts(args)
            out["renderer"] = bench_renderer(args)
    elif args.workload == "pathcuts":
        r = bench_path_cuts(args)
        out = {"metric": METRIC, "value": r["tuples_per_s"], "unit": "admissible tuples/s", "ms_per_step": r["full_ms"],
               "dtype": "f32 (directed-rounding intervals)", "config": {"workload": r["workload"]},
               "cpu_baseline": r.get("cpu_baseline"), "path_cuts": r}
    elif args.workload == "c5":
        r = bench_c5(args)
        out = {"metric": METRIC, "value": round(r["params"] / (r["backward_ms"] * 1e-3), 1),
               "unit": "height gradients/s", "ms_per_step": r["forward_ms"] + r["backward_ms"], "dtype": "f32+f64",
               "config": {"workload": r["workload"]}, "c5": r}
    elif args.workload == "renderer":
        r = bench_renderer(args)
        out = {"metric": METRIC, "value": r["frames_per_s"], "unit": "frames/s", "ms_per_step": r["ms_per_frame"],
               "dtype": "f32+f64", "config": {"workload": r["workload"]}, "renderer": r}
    elif args.workload == "c4":
        r = bench_c4(args)
        out = {"metric": METRIC, "value": r["batched_paths_per_s"], "unit": "paths/s",
               "ms_per_step": r["batched_graph_ms"], "dtype": "f32+f64",
               "config": {"workload": r["workload"]}, "c4": r}
    else:
        r = bench_lbvh(args)
        dom = max(r["stage_GBps"], key=r["stage_ms"].get)
        ach = r["stage_GBps"][dom]
        out = {"metric": METRIC, "value": r["value"], "unit": "Mtri/s", "ms_per_step": r["ms_per_step"],
               "dtype": "u32", "config": {"workload": r["workload"], "n_triangles": r["n_triangles"]},
               "roofline": {"bound": "hbm", "kernel": LBVH_KERNEL[dom], "achieved": ach, "peak": HBM_PEAK_GBPS,
                            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                            "bytes_per_unit": STAGE_BYTES_PER_TRI[dom], "units": r["n_triangles"],
                            "traffic": profiled_traffic("lbvh 16M", LBVH_KERNEL[dom])},
               "cpu_baseline": r.get("cpu_baseline"), "lbvh": r}
    out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "data": "synthetic"})
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
