"""Host-side mirror of the reference's sort and BVH builder classes.

`OnesweepSort` and `BVHGPUBuilder` keep the reference's names, argument
meaning and ownership (tinyhiponesweep.h:572-668, minimum_lbvh.h:741-900);
the work happens in libpalz.so's gfx950 kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import types as T
from ._lib import call, lib
from .device import DeviceBuffer, _sp, download_ptr


class OnesweepSort:
    """tinyhiponesweep::OnesweepSort over the C ABI."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("palz_sorter_create", C.byref(h))
        self.handle = h.value

    def reserve(self, n: int) -> None:
        call("palz_sorter_reserve", C.c_void_p(self.handle), C.c_uint32(n))

    def sort(self, words: int, tmp: int, n: int, start_bit: int, end_bit: int, stream=None,
             values: int = 0, tmp_values: int = 0) -> None:
        """OnesweepSort::sort(KeyValueSoA64{words, values}, KeyValueSoA64{tmp, tmp_values},
        n, startBit, endBit, stream): stable LSD sort of device words in place."""
        call("palz_sort_u64", C.c_void_p(self.handle), C.c_void_p(words), C.c_void_p(tmp),
             C.c_void_p(values or None), C.c_void_p(tmp_values or None), C.c_uint32(n),
             int(start_bit), int(end_bit), _sp(stream))

    def sort32(self, keys: int, tmp: int, n: int, start_bit: int, end_bit: int, stream=None,
               values: int = 0, tmp_values: int = 0) -> None:
        """OnesweepSort::sort(KeyValueSoA32{keys, values}, KeyValueSoA32{tmp, tmp_values},
        n, startBit, endBit, stream): stable LSD sort of device u32 keys in place."""
        call("palz_sort_u32", C.c_void_p(self.handle), C.c_void_p(keys), C.c_void_p(tmp),
             C.c_void_p(values or None), C.c_void_p(tmp_values or None), C.c_uint32(n),
             int(start_bit), int(end_bit), _sp(stream))

    def __del__(self):
        try:
            if self.handle:
                lib().palz_sorter_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


class BVHGPUBuilder:
    """minimum_lbvh::BVHGPUBuilder over the C ABI.

    After build(), m_internals / m_rootNode / m_sceneAABB are device pointers
    owned by the builder, as in the reference (minimum_lbvh.h:891-899).
    """

    def __init__(self):
        h = C.c_void_p()
        call("palz_lbvh_create", C.byref(h))
        self.handle = h.value
        self.m_nTriangles = 0

    def build(self, triangles: int, n_triangles: int, build_option: int, sorter: OnesweepSort,
              stream=None, morton_bits: int = 32) -> None:
        call("palz_lbvh_build", C.c_void_p(self.handle), C.c_void_p(triangles), int(n_triangles),
             C.c_uint32(build_option), int(morton_bits), C.c_void_p(sorter.handle), _sp(stream))
        self.m_nTriangles = int(n_triangles)

    def widen(self, stream=None) -> int:
        """Builds the BVH4 view (palz_wide_node per internal node) of the last
        build on `stream`; returns its device pointer."""
        w = C.c_void_p()
        call("palz_lbvh_widen", C.c_void_p(self.handle), C.byref(w), _sp(stream))
        return int(w.value or 0)

    def normal_bounds(self, attribs: int, stream=None) -> DeviceBuffer:
        """InternalNormalBound per internal node of the last build (the
        bottom-up pass of main.cpp:248-295), from the device TriangleAttrib
        array `attribs`; returns the n-1 records."""
        n_int = max(self.m_nTriangles - 1, 0)
        out = DeviceBuffer(max(n_int, 1) * T.NORMAL_BOUND.itemsize)
        call("palz_lbvh_normal_bounds", C.c_void_p(self.handle), C.c_void_p(attribs), C.c_void_p(out.ptr),
             _sp(stream))
        return out

    def set_timing(self, enable: bool = True) -> None:
        call("palz_lbvh_set_timing", C.c_void_p(self.handle), int(enable))

    STAGES = ("scene_aabb", "morton_hist", "sort_pass0", "sort_pass1", "sort_pass2",
              "sort_pass3", "build", "build_levels")

    def timings(self) -> dict:
        """Per-stage kernel durations (ms) of the last build, from HIP events."""
        ms = (C.c_float * 8)()
        k = C.c_int(0)
        call("palz_lbvh_timings", C.c_void_p(self.handle), ms, 8, C.byref(k))
        return {self.STAGES[i]: float(ms[i]) for i in range(k.value)}

    def _result(self):
        nodes, root, aabb = C.c_void_p(), C.c_void_p(), C.c_void_p()
        n = C.c_int(0)
        call("palz_lbvh_result", C.c_void_p(self.handle), C.byref(nodes), C.byref(root),
             C.byref(aabb), C.byref(n))
        return nodes.value or 0, root.value or 0, aabb.value or 0, n.value

    @property
    def m_internals(self) -> int:
        return self._result()[0]

    @property
    def m_rootNode(self) -> int:
        return self._result()[1]

    @property
    def m_sceneAABB(self) -> int:
        return self._result()[2]

    def empty(self) -> bool:
        return self.m_nTriangles == 0

    # -- host copies for checks ------------------------------------------------
    def download_nodes(self) -> np.ndarray:
        nodes, _, _, n = self._result()
        return download_ptr(nodes, T.INTERNAL_NODE, max(n - 1, 0))

    def download_root(self) -> int:
        _, root, _, _ = self._result()
        return int(download_ptr(root, np.uint32, 1)[0])

    def download_scene_aabb(self) -> np.ndarray:
        _, _, aabb, _ = self._result()
        return download_ptr(aabb, T.AABB, 1)[0]

    def download_sorted(self):
        w, d = C.c_void_p(), C.c_void_p()
        call("palz_lbvh_sorted", C.c_void_p(self.handle), C.byref(w), C.byref(d))
        n = self.m_nTriangles
        return (download_ptr(w.value, np.uint64, n),
                download_ptr(d.value, np.uint8, max(n - 1, 0)))

    def __del__(self):
        try:
            if self.handle:
                lib().palz_lbvh_destroy(C.c_void_p(self.handle))
                self.handle = None
        except Exception:
            pass


def build_lbvh_from_host(tris: np.ndarray, morton_bits: int = 32, build_option: int = 0,
                         sorter: OnesweepSort | None = None):
    """Upload triangles (N x 9 float32 or TRIANGLE records) and build; returns
    (builder, device triangle buffer, sorter)."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    dt = DeviceBuffer.from_array(tris)
    sorter = sorter or OnesweepSort()
    b = BVHGPUBuilder()
    b.build(dt.ptr, tris.shape[0], build_option, sorter, morton_bits=morton_bits)
    return b, dt, sorter
