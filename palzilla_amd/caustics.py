"""Host-side mirror of the reference's caustic step (PKRenderer::loadFrame and
PKRenderer::step, pk.h:1246-1627) over the C ABI.

A `GpuScene` owns the device copies of one CausticScene and its LBVH.  The
methods follow the reference's order: photonTrace -> PathCache (here the
deterministic tuple table) -> lookupFlatten / sort / reorder per event
descriptor -> solveSpecularPath_K.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import types as T
from ._lib import call, lib
from .device import DeviceBuffer, _sp, download_ptr
from .lbvh import BVHGPUBuilder, OnesweepSort
from .scenes import FIRST_DIFFUSE, CausticScene


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class SceneView(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("attribs", C.c_void_p), ("nodes", C.c_void_p), ("root", C.c_void_p),
                ("n_tris", C.c_uint32), ("pad_", C.c_uint32)]


class PhotonParams(C.Structure):
    _fields_ = [("p_light", Float3), ("eta_min", C.c_float), ("eta_max", C.c_float), ("iteration", C.c_int32),
                ("min_throughput", C.c_float), ("photons_per_triangle", C.c_uint32)]


class Camera(C.Structure):
    """palz_camera: RayGenerator's origin and image-plane half axes (camera.h:4-41)."""
    _fields_ = [("origin", Float3), ("right", Float3), ("up", Float3)]

    @staticmethod
    def lookat(eye, center, up, fovy: float, width: int, height: int) -> "Camera":
        """RayGenerator::lookat (camera.h:10-22), in FP32."""
        f32 = np.float32
        eye, center, up = (np.asarray(v, dtype=f32) for v in (eye, center, up))

        def normalize(v):
            return (v * (f32(1.0) / np.sqrt(f32(np.dot(v, v))))).astype(f32)

        f = normalize(center - eye)
        s = normalize(np.cross(f, up).astype(f32))
        u = np.cross(s, f).astype(f32)
        tan_y = f32(np.tan(f32(fovy) * f32(0.5)))
        tan_x = f32(tan_y / f32(height) * f32(width))
        return Camera(Float3(*eye), Float3(*(s * tan_x)), Float3(*(u * tan_y)))


class PrimaryParams(C.Structure):
    """palz_primary_params (what solvePrimary takes besides the scene)."""
    _fields_ = [("p_light", Float3), ("light_intensity", C.c_float), ("radiance_clamp", C.c_float),
                ("cauchy_a", C.c_float), ("cauchy_b", C.c_float), ("iteration", C.c_int32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("use_lens", C.c_uint32), ("lens_distance", C.c_float),
                ("lens_thickness", C.c_float), ("lens_R", C.c_float), ("albedo", Float3), ("texture", C.c_void_p),
                ("texture_width", C.c_uint32), ("texture_height", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("attribs", C.c_void_p), ("nodes", C.c_void_p), ("root", C.c_void_p),
                ("points", C.c_void_p), ("lambdas", C.c_void_p), ("image", C.c_void_p),
                ("n_points", C.c_uint32), ("n_lambdas", C.c_uint32), ("p_light", Float3),
                ("light_intensity", C.c_float), ("radiance_clamp", C.c_float), ("cauchy_a", C.c_float),
                ("cauchy_b", C.c_float), ("wide", C.c_void_p), ("image_fixed", C.c_void_p)]


assert C.sizeof(SceneView) == 40
assert C.sizeof(PhotonParams) == 32
assert C.sizeof(SceneDesc) == 112

PHOTON_RECORD = np.dtype([("photon", "<u4"), ("event", "<u4"), ("tris", "<i4", 4), ("params", "<f4", 8),
                          ("p_final", "<f4", 3), ("throughput", "<f4")])
assert PHOTON_RECORD.itemsize == 72
PATH = np.dtype([("scene", "<u4"), ("pixel", "<u4"), ("event", "<u4"), ("photon", "<u4"), ("tris", "<i4", 4),
                 ("params", "<f4", 8)])
assert PATH.itemsize == 64
PATH_RESULT = np.dtype([("status", "<u4"), ("iterations", "<i4"), ("contribution", "<f4"), ("params", "<f4", 8),
                        ("footprint", "<f4")])
assert PATH_RESULT.itemsize == 48

PATH_NOT_CONVERGED, PATH_CONVERGED, PATH_CONTRIBUTES, PATH_CULLED = 0, 1, 2, 3   # palz_path_result.status
REFINE_NONE, REFINE_VERTICES, REFINE_ALL = 0, 1, 2                # palz_batched_solve refine modes

# Arithmetic of one Newton iteration of the reference solver (solveConstraints,
# pk.h:477-590: 2K single-tangent residual/Jacobian passes with the cost, then
# the Householder QR step, the update and the triangle test), counted exactly
# with a FLOP-counting scalar (+ - * / sqrt = 1).  A residual evaluation costs
# EVAL_BASE[K] + EVAL_PER_T[K] * (number of refractive vertices); the step after
# a non-converged evaluation costs at most STEP[K].  These are the algorithmic
# FLOPs of the solver roofline (the kernels here share the primal chain across
# tangents and execute fewer).
ITERATION_FLOPS_EVAL_BASE = {1: 591, 2: 2322, 3: 5193}
ITERATION_FLOPS_EVAL_PER_T = {1: 56, 2: 112, 3: 168}
ITERATION_FLOPS_STEP = {1: 85, 2: 361, 3: 941}


def iteration_flops(event: int) -> tuple[int, int]:
    """(evaluation, step) FLOPs of one reference Newton iteration for an event."""
    K = event >> 24
    n_t = bin(event & ((1 << K) - 1)).count("1")
    return ITERATION_FLOPS_EVAL_BASE[K] + ITERATION_FLOPS_EVAL_PER_T[K] * n_t, ITERATION_FLOPS_STEP[K]


LIGHT_INTENSITY = 16.0   # m_lightIntencity (pk.h:1686)
RADIANCE_CLAMP = 12.0    # m_radianceClamp (pk.h:1688)
MIN_THROUGHPUT = 0.05    # m_minThroughput (pk.h:1687)
CELL_STEP = 0.01         # m_pathCache.init(0.01f) (pk.h:1376)


def event_u32(events: str) -> int:
    """EventDescriptor::asU32 for a string of R/T events, e.g. "TT"."""
    bits = 0
    for i, e in enumerate(events):
        if e == "T":
            bits |= 1 << i
    return (len(events) << 24) | bits


def event_str(ev: int) -> str:
    return "".join("T" if (ev >> i) & 1 else "R" for i in range(ev >> 24))


class TupleTable:
    def __init__(self):
        h = C.c_void_p()
        call("palz_tuple_table_create", C.byref(h))
        self.handle = h.value

    def build(self, records: int, n_records: int, step: float, sorter: OnesweepSort, stream=None):
        call("palz_tuple_table_build", C.c_void_p(self.handle), C.c_void_p(records), C.c_uint32(n_records),
             C.c_float(step), C.c_void_p(sorter.handle), _sp(stream))

    def info(self) -> dict:
        v = [C.c_uint32(0) for _ in range(4)]
        call("palz_tuple_table_info", C.c_void_p(self.handle), *[C.byref(x) for x in v])
        return dict(zip(("n_valid", "n_tuples", "n_entries", "n_cells"), (x.value for x in v)))

    def __del__(self):
        try:
            if self.handle:
                lib().palz_tuple_table_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


class GpuScene:
    """Device copies of a CausticScene plus its LBVH and working buffers."""

    def __init__(self, scene: CausticScene, sorter: OnesweepSort | None = None, stream=None):
        self.scene = scene
        self.stream = stream
        self.sorter = sorter or OnesweepSort()
        self.d_tris = DeviceBuffer.from_array(scene.tris)
        self.d_attribs = DeviceBuffer.from_array(scene.attribs)
        self.d_points = DeviceBuffer.from_array(scene.points)
        self.d_lambdas = DeviceBuffer.from_array(scene.lambdas.astype(np.float32))
        self.n_points = scene.points.shape[0]
        self.n_lambdas = scene.lambdas.shape[0]
        self.d_image = DeviceBuffer(4 * self.n_points * self.n_lambdas)
        self.d_image.zero(stream)
        self.bvh = BVHGPUBuilder()
        self.bvh.build(self.d_tris.ptr, scene.n_tris, 0, self.sorter, stream=stream, morton_bits=32)
        self.wide = self.bvh.widen(stream)  # BVH4 view for the batched solver's occlusion rays
        self.table = TupleTable()
        self.d_records = None
        self.d_status = DeviceBuffer(16)
        self.d_status.zero(stream)
        self.d_count = DeviceBuffer(16)

    # -- C ABI records ---------------------------------------------------------
    def view(self) -> SceneView:
        return SceneView(self.d_tris.ptr, self.d_attribs.ptr, self.bvh.m_internals, self.bvh.m_rootNode,
                         self.scene.n_tris, 0)

    def desc(self, image_ptr: int | None = None, wide: bool = True, image_fixed: int | None = None) -> SceneDesc:
        """The scene's palz_scene_desc; wide=False leaves out the BVH4 view;
        image_fixed names a 32.32 fixed-point image to sum into instead."""
        s = self.scene
        return SceneDesc(self.d_tris.ptr, self.d_attribs.ptr, self.bvh.m_internals, self.bvh.m_rootNode,
                         self.d_points.ptr, self.d_lambdas.ptr, image_ptr or self.d_image.ptr, self.n_points,
                         self.n_lambdas, Float3(*s.p_light), LIGHT_INTENSITY, RADIANCE_CLAMP, s.cauchy_a, s.cauchy_b,
                         self.wide if wide else None, image_fixed)

    def photon_params(self, iteration: int = 0) -> PhotonParams:
        s = self.scene
        return PhotonParams(Float3(*s.p_light), s.eta_min, s.eta_max, iteration, MIN_THROUGHPUT,
                            s.photons_per_triangle)

    @property
    def n_photon_slots(self) -> int:
        return self.scene.n_tris * self.scene.photons_per_triangle

    # -- pipeline --------------------------------------------------------------
    def trace(self, iteration: int = 0) -> None:
        n = self.n_photon_slots
        if self.d_records is None or self.d_records.nbytes < n * PHOTON_RECORD.itemsize:
            self.d_records = DeviceBuffer(n * PHOTON_RECORD.itemsize)
        v, p = self.view(), self.photon_params(iteration)
        call("palz_photon_trace", C.byref(v), C.byref(p), C.c_void_p(self.d_records.ptr),
             C.c_void_p(self.d_status.ptr), _sp(self.stream))

    def trace_status(self) -> int:
        return int(self.d_status.download(np.uint32, 1, self.stream)[0])

    def records(self) -> np.ndarray:
        return self.d_records.download(PHOTON_RECORD, self.n_photon_slots, self.stream)

    def build_table(self, step: float = CELL_STEP) -> dict:
        self.table.build(self.d_records.ptr, self.n_photon_slots, step, self.sorter, self.stream)
        return self.table.info()

    def lookup(self, event: int, capacity: int, d_paths: DeviceBuffer | None = None, scene_index: int = 0):
        """lookupFlatten for one event (0 = all); returns (paths buffer, count).
        Reads the count back to the host, as the reference does (pk.h:1502-1503)."""
        if d_paths is None or d_paths.nbytes < capacity * PATH.itemsize:
            d_paths = DeviceBuffer(max(capacity, 1) * PATH.itemsize)
        call("palz_lookup_flatten", C.c_void_p(self.table.handle), C.c_uint32(event), C.c_void_p(self.d_points.ptr),
             C.c_uint32(self.n_points), C.c_uint32(scene_index), C.c_void_p(d_paths.ptr), C.c_uint32(capacity),
             C.c_void_p(self.d_count.ptr), C.c_void_p(self.sorter.handle), _sp(self.stream))
        n = int(self.d_count.download(np.uint32, 1, self.stream)[0])
        return d_paths, n

    def solve_plain(self, K: int, event: int, d_paths: DeviceBuffer, n_paths: int,
                    d_results: DeviceBuffer | None = None, image_fixed: int | None = None) -> None:
        d = self.desc(image_fixed=image_fixed)
        call("palz_solve_specular", int(K), C.c_uint32(event), C.byref(d), C.c_void_p(d_paths.ptr),
             C.c_uint32(n_paths), C.c_void_p(None), C.c_void_p(d_results.ptr if d_results else None),
             _sp(self.stream))

    def intersect_rays(self, ro: np.ndarray, rd: np.ndarray, query: int = 0,
                       t_max: float = 3.402823466e38) -> np.ndarray:
        """palz_intersect_rays over host rays (n x 3 origins and directions):
        query 0 stackfull closest hit, 1 stackfree closest hit, 2 any hit;
        returns the palz_hit records."""
        from .types import HIT
        ro = np.ascontiguousarray(ro, np.float32).reshape(-1, 3)
        rd = np.ascontiguousarray(rd, np.float32).reshape(-1, 3)
        n = ro.shape[0]
        d_ro, d_rd = DeviceBuffer.from_array(ro), DeviceBuffer.from_array(rd)
        d_hits = DeviceBuffer(max(n, 1) * HIT.itemsize)
        v = self.view()
        call("palz_intersect_rays", C.byref(v), C.c_void_p(d_ro.ptr), C.c_void_p(d_rd.ptr), C.c_uint32(n),
             int(query), C.c_float(t_max), C.c_void_p(d_hits.ptr), _sp(self.stream))
        return d_hits.download(HIT, n, self.stream)

    def image(self) -> np.ndarray:
        return self.d_image.download(np.float32, self.n_points * self.n_lambdas, self.stream).reshape(
            self.n_lambdas, self.scene.height, self.scene.width)

    def clear_image(self) -> None:
        self.d_image.zero(self.stream)


def image_fixed_to_float(d_fixed: DeviceBuffer, n: int, stream=None) -> np.ndarray:
    """palz_image_fixed_to_float of n entries, downloaded."""
    out = DeviceBuffer(max(n, 1) * 4)
    call("palz_image_fixed_to_float", C.c_void_p(d_fixed.ptr), C.c_void_p(out.ptr), C.c_size_t(n), _sp(stream))
    return out.download(np.float32, n, stream)


def download_paths(d_paths: DeviceBuffer, n: int) -> np.ndarray:
    return d_paths.download(PATH, n)


def download_paths_at(ptr: int, start: int, n: int) -> np.ndarray:
    """Paths [start, start + n) of a device path list."""
    return download_ptr(ptr + start * PATH.itemsize, PATH, n)


def download_results(d_results: DeviceBuffer, n: int) -> np.ndarray:
    return d_results.download(PATH_RESULT, n)


# -- batched solver --------------------------------------------------------------------------------

class LaunchConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("chunk", C.c_int32),
                ("finish_block", C.c_int32)]


STAT_ROWS = ("items", "converged", "contributing", "iterations")


class BatchedSolver:
    """palz_batched_*: chain-length partition, persistent Newton kernels,
    finish kernels with the FP64 refinement step, autotuner and hipGraph.

    Results are indexed by (length-sorted path, wavelength); `order()` maps a
    sorted path back to its index in the list given to `prepare`."""

    def __init__(self, sorter: OnesweepSort, stream=None):
        h = C.c_void_p()
        call("palz_batched_create", C.byref(h))
        self.handle = h.value
        self.sorter = sorter
        self.stream = stream
        self.n_paths = 0

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().palz_batched_destroy(C.c_void_p(self.handle))
            except Exception:
                pass
            self.handle = None

    def prepare(self, d_paths: DeviceBuffer | int, n_paths: int, d_count: int | None = None) -> None:
        ptr = d_paths.ptr if isinstance(d_paths, DeviceBuffer) else d_paths
        call("palz_batched_prepare", C.c_void_p(self.handle), C.c_void_p(ptr), C.c_uint32(n_paths),
             C.c_void_p(d_count), C.c_void_p(self.sorter.handle), _sp(self.stream))
        self.n_paths = n_paths

    def pointers(self) -> tuple[int, int, int]:
        o, p, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("palz_batched_order", C.c_void_p(self.handle), C.byref(o), C.byref(p), C.byref(r))
        return o.value, p.value, r.value

    def order(self) -> np.ndarray:
        o, _, _ = self.pointers()
        return download_ptr(o, np.uint32, self.ranges()[5], self.stream)

    def sorted_paths(self) -> np.ndarray:
        _, p, _ = self.pointers()
        return download_ptr(p, PATH, self.ranges()[5], self.stream)

    def ranges(self) -> np.ndarray:
        """[count with chain length <= k for k = 0..4, path count]."""
        _, _, r = self.pointers()
        return download_ptr(r, np.uint32, 6, self.stream)

    def set_config(self, K: int, block: int, blocks_per_cu: int, chunk: int, finish_block: int) -> None:
        c = LaunchConfig(block, blocks_per_cu, chunk, finish_block)
        call("palz_batched_set_config", C.c_void_p(self.handle), int(K), C.byref(c))

    def get_config(self, K: int) -> dict:
        c = LaunchConfig()
        call("palz_batched_get_config", C.c_void_p(self.handle), int(K), C.byref(c))
        return {f: getattr(c, f) for f, _ in LaunchConfig._fields_}

    def solve(self, d_scenes: DeviceBuffer, n_scenes: int, n_lambdas: int, refine: int = 1,
              d_results: DeviceBuffer | None = None, d_refined: DeviceBuffer | None = None,
              graph: bool = False) -> None:
        name = "palz_batched_solve_graph" if graph else "palz_batched_solve"
        call(name, C.c_void_p(self.handle), C.c_void_p(d_scenes.ptr), C.c_uint32(n_scenes), C.c_uint32(n_lambdas),
             int(refine), C.c_void_p(d_results.ptr if d_results else None),
             C.c_void_p(d_refined.ptr if d_refined else None), _sp(self.stream))

    TIMED_KERNELS = ("newton_k1", "newton_k2", "newton_k3", "compact", "occlude_k1", "finish_k1", "occlude_k2",
                     "finish_k2", "occlude_k3", "finish_k3", "refine", "cull")

    def set_timing(self, enable: bool) -> None:
        call("palz_batched_set_timing", C.c_void_p(self.handle), int(bool(enable)))

    def timings(self) -> dict:
        ms = (C.c_float * len(self.TIMED_KERNELS))()
        call("palz_batched_timings", C.c_void_p(self.handle), ms)
        return dict(zip(self.TIMED_KERNELS, (float(v) for v in ms)))

    def set_culling(self, mode: int | bool) -> None:
        """0/False: off; 1/True: interval test; 2: interval test, then affine test."""
        mode = int(mode)
        call("palz_batched_set_culling", C.c_void_p(self.handle), mode)

    def stats(self) -> dict:
        s = (C.c_uint64 * 16)()
        call("palz_batched_stats", C.c_void_p(self.handle), s, _sp(self.stream))
        a = np.frombuffer(s, dtype=np.uint64).copy()
        out = {f"K{k + 1}": dict(zip(STAT_ROWS, (int(v) for v in a[4 * k:4 * k + 4]))) for k in range(3)}
        out["skipped"] = int(a[15])
        out["culled"] = int(a[12])
        out["bad_scene"] = int(a[14])
        return out

    def set_fused(self, mode: int | bool) -> None:
        """A bit mask: 1 = one persistent Newton kernel over every chain
        length, 2 = one occlusion kernel over every chain length, 3 (or True)
        = both, 0 (or False) = one of each per chain length; -1 (the
        default): chosen by the autotuner (per chain length until it has
        run)."""
        mode = 3 if mode is True else int(mode)
        call("palz_batched_set_fused", C.c_void_p(self.handle), mode)

    def get_fused(self) -> tuple[int, int]:
        """(mode set, the mask solves now run with)."""
        mode, active = C.c_int(), C.c_int()
        call("palz_batched_get_fused", C.c_void_p(self.handle), C.byref(mode), C.byref(active))
        return mode.value, active.value

    def set_tune_cache(self, path: str | None) -> None:
        """The autotuner's tile-shape cache file (None: no cache)."""
        call("palz_batched_set_tune_cache", C.c_void_p(self.handle),
             None if path is None else os.fsencode(path))

    def set_branches(self, mode: int) -> None:
        """Captured solves as three graph branches (1), one stream (0) or
        decided from the chain-length mix (-1, the default)."""
        call("palz_batched_set_branches", C.c_void_p(self.handle), int(mode))

    def path_stats(self) -> dict:
        """Per chain length: paths of the last solve with a converged item."""
        s = (C.c_uint64 * 3)()
        call("palz_batched_path_stats", C.c_void_p(self.handle), s, _sp(self.stream))
        return {f"K{k + 1}": int(s[k]) for k in range(3)}

    def autotune(self, d_scenes: DeviceBuffer, n_scenes: int, n_lambdas: int, refine: int = 1) -> list[dict]:
        rep = (C.c_int32 * 18)()
        call("palz_batched_autotune", C.c_void_p(self.handle), C.c_void_p(d_scenes.ptr), C.c_uint32(n_scenes),
             C.c_uint32(n_lambdas), int(refine), rep, _sp(self.stream))
        keys = ("block", "blocks_per_cu", "chunk", "finish_block", "newton_us", "finish_us")
        return [dict(zip(keys, (int(v) for v in rep[6 * k:6 * k + 6]))) for k in range(3)]


def solve_primary(g: "GpuScene", camera: Camera, pp: PrimaryParams, d_out: DeviceBuffer | None = None,
                  d_accum: DeviceBuffer | None = None, stream=None) -> tuple[DeviceBuffer, DeviceBuffer]:
    """palz_solve_primary over the scene `g`: (FirstDiffuse per pixel,
    float4 accumulators).  Fresh buffers start at zero."""
    n = int(pp.width) * int(pp.height)
    if d_out is None:
        d_out = DeviceBuffer(max(n, 1) * FIRST_DIFFUSE.itemsize)
        d_out.zero(stream)
    if d_accum is None:
        d_accum = DeviceBuffer(max(n, 1) * 16)
        d_accum.zero(stream)
    v = g.view()
    call("palz_solve_primary", C.byref(v), C.byref(camera), C.byref(pp), C.c_void_p(d_out.ptr),
         C.c_void_p(d_accum.ptr), _sp(stream))
    return d_out, d_accum


def spectral_resolve(d_images: DeviceBuffer | int, d_lambdas: DeviceBuffer | int, d_weights: DeviceBuffer | int,
                     d_points: DeviceBuffer | int, n_points: int, n_lambdas: int, n_frames: float = 1.0,
                     stream=None) -> tuple[DeviceBuffer, DeviceBuffer]:
    """palz_spectral_resolve: (accumulators float4 x n_points, RGBA8 pixels) of
    the per-wavelength irradiance images (pk.h:960-1031, main_gpu.cu:499-513)."""
    ptr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else int(b)  # noqa: E731
    acc = DeviceBuffer(16 * max(n_points, 1))
    pix = DeviceBuffer(4 * max(n_points, 1))
    call("palz_spectral_resolve", C.c_void_p(ptr(d_images)), C.c_void_p(ptr(d_lambdas)),
         C.c_void_p(ptr(d_weights)), C.c_void_p(ptr(d_points)), C.c_uint32(n_points), C.c_uint32(n_lambdas),
         C.c_float(n_frames), C.c_void_p(acc.ptr), C.c_void_p(pix.ptr), _sp(stream))
    return acc, pix


class PathCuts:
    """palz_path_cuts_*: node-level path cuts (traverseAdmissibleNodes,
    pk.h:1760-2212) from the light to every receiver point of a scene,
    emitting palz_path records that the solvers take as they are."""

    def __init__(self, stream=None):
        h = C.c_void_p()
        call("palz_path_cuts_create", C.byref(h))
        self.handle = h.value
        self.stream = stream
        self.d_count = DeviceBuffer(16)
        self.d_status = DeviceBuffer(16)
        self.d_status.zero(stream)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().palz_path_cuts_destroy(C.c_void_p(self.handle))
            except Exception:
                pass
            self.handle = None

    def enumerate(self, desc: SceneDesc, d_normal_bounds: DeviceBuffer | int, event: int, eta_lo: float,
                  eta_hi: float, capacity: int, d_paths: DeviceBuffer | None = None,
                  scene_index: int = 0) -> tuple[DeviceBuffer | None, int]:
        """(path buffer, total count); the buffer holds min(total, capacity) paths."""
        if d_paths is None and capacity:
            d_paths = DeviceBuffer(capacity * PATH.itemsize)
        nb = d_normal_bounds.ptr if isinstance(d_normal_bounds, DeviceBuffer) else int(d_normal_bounds)
        call("palz_path_cuts_enumerate", C.c_void_p(self.handle), C.byref(desc), C.c_void_p(nb), C.c_uint32(event),
             C.c_float(eta_lo), C.c_float(eta_hi), C.c_uint32(scene_index),
             C.c_void_p(d_paths.ptr if d_paths else None), C.c_uint32(capacity), C.c_void_p(self.d_count.ptr),
             C.c_void_p(self.d_status.ptr), _sp(self.stream))
        n = int(self.d_count.download(np.uint32, 1, self.stream)[0])
        return d_paths, n

    def status(self) -> int:
        return int(self.d_status.download(np.uint32, 1, self.stream)[0])


class RendererParams(C.Structure):
    """palz_renderer_params."""
    _fields_ = [("p_light", Float3), ("light_intensity", C.c_float), ("radiance_clamp", C.c_float),
                ("min_throughput", C.c_float), ("cauchy_a", C.c_float), ("cauchy_b", C.c_float),
                ("photons_per_triangle", C.c_uint32), ("use_lens", C.c_uint32), ("lens_distance", C.c_float),
                ("lens_thickness", C.c_float), ("lens_R", C.c_float), ("albedo", Float3),
                ("path_capacity", C.c_uint32), ("refine", C.c_int32), ("solve_tttt", C.c_uint32)]


assert C.sizeof(RendererParams) == 76


class Renderer:
    """palz_renderer_*: the library's PKRenderer-shaped host driver (C++)."""

    def __init__(self, width: int, height: int, lambdas, weights, stream=None):
        lam = np.ascontiguousarray(lambdas, dtype=np.float32)
        w = np.ascontiguousarray(weights, dtype=np.float32)
        h = C.c_void_p()
        call("palz_renderer_create", int(width), int(height), lam.ctypes.data_as(C.c_void_p),
             w.ctypes.data_as(C.c_void_p), lam.shape[0], C.byref(h))
        self.handle = h.value
        self.width, self.height = width, height
        self.stream = stream

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().palz_renderer_destroy(C.c_void_p(self.handle))
            except Exception:
                pass
            self.handle = None

    def params(self) -> RendererParams:
        p = RendererParams()
        call("palz_renderer_get_params", C.c_void_p(self.handle), C.byref(p))
        return p

    def set_params(self, p: RendererParams) -> None:
        call("palz_renderer_set_params", C.c_void_p(self.handle), C.byref(p))

    def set_camera(self, cam: Camera) -> None:
        call("palz_renderer_set_camera", C.c_void_p(self.handle), C.byref(cam))

    def load_scene(self, tris: np.ndarray, attribs: np.ndarray) -> None:
        t = np.ascontiguousarray(tris, dtype=np.float32)
        a = np.ascontiguousarray(attribs)
        call("palz_renderer_load_scene", C.c_void_p(self.handle), t.ctypes.data_as(C.c_void_p),
             a.ctypes.data_as(C.c_void_p), C.c_uint32(t.shape[0]), _sp(self.stream))

    def load_obj(self, path: str, default_material: int = 0) -> None:
        call("palz_renderer_load_obj", C.c_void_p(self.handle), os.fsencode(path), int(default_material),
             _sp(self.stream))

    def step(self) -> None:
        call("palz_renderer_step", C.c_void_p(self.handle), _sp(self.stream))

    def clear(self) -> None:
        call("palz_renderer_clear", C.c_void_p(self.handle), _sp(self.stream))

    def resolve(self) -> np.ndarray:
        out = np.zeros(self.width * self.height, dtype=np.uint32)
        call("palz_renderer_resolve", C.c_void_p(self.handle), out.ctypes.data_as(C.c_void_p), _sp(self.stream))
        return out

    def buffers(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(first-diffuse records, accumulators (n, 4), counts [status, paths, T,T,T,T paths])."""
        pts, acc, cnt = C.c_void_p(), C.c_void_p(), C.c_void_p()
        call("palz_renderer_buffers", C.c_void_p(self.handle), C.byref(pts), C.byref(acc), C.byref(cnt))
        n = self.width * self.height
        return (download_ptr(pts.value, FIRST_DIFFUSE, n, self.stream),
                download_ptr(acc.value, np.float32, 4 * n, self.stream).reshape(n, 4),
                download_ptr(cnt.value, np.uint32, 3, self.stream))


def eta_range(scene: CausticScene) -> tuple[float, float]:
    """The refractive indices of the scene's wavelengths as the solvers
    compute them (float32 Cauchy), as (lowest, highest)."""
    lam = scene.lambdas.astype(np.float32)
    eta = np.float32(scene.cauchy_a) + np.float32(scene.cauchy_b) / (lam * lam)
    return float(eta.min()), float(eta.max())


def upload_scene_descs(descs: list[SceneDesc], stream=None) -> DeviceBuffer:
    arr = (SceneDesc * len(descs))(*descs)
    raw = np.frombuffer(bytes(arr), dtype=np.uint8)
    return DeviceBuffer.from_array(raw, stream)


class SceneBatch:
    """Many scenes in one path list: each scene is traced, its tuple table
    built and looked up over all its receiver points (scene index = position
    in the batch), and the path lists are concatenated on the device."""

    def __init__(self, scene_list: list[CausticScene], sorter: OnesweepSort | None = None, stream=None):
        self.sorter = sorter or OnesweepSort()
        self.stream = stream
        self.scenes = [GpuScene(s, self.sorter, stream) for s in scene_list]
        L = {g.n_lambdas for g in self.scenes}
        if len(L) != 1:
            raise ValueError("all scenes of a batch need the same number of wavelengths")
        self.n_lambdas = L.pop()
        self.counts: list[int] = []
        self.offsets: list[int] = []
        self.d_paths: DeviceBuffer | None = None
        self.n_paths = 0
        self.d_scenes: DeviceBuffer | None = None

    def build(self, iteration: int = 0, capacity_per_scene: int = 1 << 22) -> int:
        lists = []
        for i, g in enumerate(self.scenes):
            g.trace(iteration)
            g.build_table()
            d, n = g.lookup(0, capacity=capacity_per_scene, scene_index=i)
            if n > capacity_per_scene:
                d, n = g.lookup(0, capacity=n, scene_index=i)
            lists.append((d, n))
        self.counts = [n for _, n in lists]
        self.offsets = list(np.cumsum([0] + self.counts[:-1]))
        self.n_paths = int(sum(self.counts))
        self.d_paths = DeviceBuffer(max(self.n_paths, 1) * PATH.itemsize)
        for (d, n), off in zip(lists, self.offsets):
            if n:
                call("palz_memcpy_d2d", C.c_void_p(self.d_paths.ptr + int(off) * PATH.itemsize), C.c_void_p(d.ptr),
                     C.c_size_t(n * PATH.itemsize), _sp(self.stream))
        self.d_scenes = upload_scene_descs([g.desc() for g in self.scenes], self.stream)
        return self.n_paths

    def clear_images(self) -> None:
        for g in self.scenes:
            g.clear_image()


def interval_cull(d_scenes: DeviceBuffer, n_scenes: int, d_paths: DeviceBuffer | int, n_paths: int,
                  stream=None) -> np.ndarray:
    """palz_interval_cull: per-path admissibility flags (1 = keep)."""
    ptr = d_paths.ptr if isinstance(d_paths, DeviceBuffer) else d_paths
    d_out = DeviceBuffer(max(n_paths, 1))
    call("palz_interval_cull", C.c_void_p(d_scenes.ptr), C.c_uint32(n_scenes), C.c_void_p(ptr), C.c_uint32(n_paths),
         C.c_void_p(None), C.c_void_p(d_out.ptr), _sp(stream))
    return d_out.download(np.uint8, n_paths, stream)


def affine_cull(d_scenes: DeviceBuffer, n_scenes: int, d_paths: DeviceBuffer | int, n_paths: int,
                stream=None) -> np.ndarray:
    """palz_affine_cull: the interval test, then the affine test on what it keeps (1 = keep)."""
    ptr = d_paths.ptr if isinstance(d_paths, DeviceBuffer) else d_paths
    d_out = DeviceBuffer(max(n_paths, 1))
    call("palz_affine_cull", C.c_void_p(d_scenes.ptr), C.c_uint32(n_scenes), C.c_void_p(ptr), C.c_uint32(n_paths),
         C.c_void_p(None), C.c_void_p(d_out.ptr), _sp(stream))
    return d_out.download(np.uint8, n_paths, stream)


class HeightParamStruct(C.Structure):
    _fields_ = [("tri_vertex", C.c_void_p), ("normal_nb", C.c_void_p), ("normal_coef", C.c_void_p),
                ("n_params", C.c_uint32), ("pad_", C.c_uint32)]


class HeightParams:
    """Device copy of a heightfield parameterisation (scenes.height_param_tables)."""

    def __init__(self, tri_vertex: np.ndarray, normal_nb: np.ndarray, normal_coef: np.ndarray):
        self.d_tv = DeviceBuffer.from_array(np.ascontiguousarray(tri_vertex, dtype=np.int32))
        self.d_nb = DeviceBuffer.from_array(np.ascontiguousarray(normal_nb, dtype=np.int32))
        self.d_coef = DeviceBuffer.from_array(np.ascontiguousarray(normal_coef, dtype=np.float32))
        self.n_params = normal_nb.shape[0]
        self.struct = HeightParamStruct(self.d_tv.ptr, self.d_nb.ptr, self.d_coef.ptr, self.n_params, 0)


def adjoint_irradiance(g: "GpuScene", hp: HeightParams, d_paths: int, n_paths: int, d_results: int | None,
                       d_weight: DeviceBuffer, d_grad: DeviceBuffer, stream=None) -> None:
    """palz_adjoint_irradiance: d_grad (n_params doubles) += dL/dh for the
    items of d_paths given dL/dI in d_weight (n_lambdas x n_points floats)."""
    d = g.desc()
    call("palz_adjoint_irradiance", C.byref(d), C.byref(hp.struct), C.c_void_p(d_paths), C.c_uint32(n_paths),
         C.c_void_p(d_results), C.c_void_p(d_weight.ptr), C.c_void_p(d_grad.ptr), _sp(stream))
