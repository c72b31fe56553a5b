"""ctypes binding of libpalz.so (the C ABI declared in include/palz.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpalz.so")

c_void_pp = C.POINTER(C.c_void_p)


class PalzError(RuntimeError):
    """A libpalz entry point returned a non-zero status."""

    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} -> status {code}: {msg}")
        self.code = code


# name -> (restype, argtypes); every function returns an int status unless noted
_SIGS = {
    "palz_last_error": (C.c_char_p, []),
    "palz_version": (C.c_int, []),
    "palz_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "palz_set_device": (C.c_int, [C.c_int]),
    "palz_malloc": (C.c_int, [c_void_pp, C.c_size_t]),
    "palz_free": (C.c_int, [C.c_void_p]),
    "palz_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "palz_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "palz_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "palz_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]),
    "palz_stream_create": (C.c_int, [c_void_pp]),
    "palz_stream_destroy": (C.c_int, [C.c_void_p]),
    "palz_stream_sync": (C.c_int, [C.c_void_p]),
    "palz_device_sync": (C.c_int, []),
    "palz_event_create": (C.c_int, [c_void_pp]),
    "palz_event_destroy": (C.c_int, [C.c_void_p]),
    "palz_event_record": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
    "palz_copy_bytes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "palz_sorter_create": (C.c_int, [c_void_pp]),
    "palz_sorter_destroy": (C.c_int, [C.c_void_p]),
    "palz_sorter_reserve": (C.c_int, [C.c_void_p, C.c_uint32]),
    "palz_sort_u64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "palz_sort_u32": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "palz_sort_u64_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                        C.c_int, C.c_void_p]),
    "palz_lbvh_create": (C.c_int, [c_void_pp]),
    "palz_lbvh_destroy": (C.c_int, [C.c_void_p]),
    "palz_lbvh_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_int,
                                  C.c_void_p, C.c_void_p]),
    "palz_lbvh_result": (C.c_int, [C.c_void_p, c_void_pp, c_void_pp, c_void_pp,
                                   C.POINTER(C.c_int)]),
    "palz_lbvh_sorted": (C.c_int, [C.c_void_p, c_void_pp, c_void_pp]),
    "palz_lbvh_widen": (C.c_int, [C.c_void_p, c_void_pp, C.c_void_p]),
    "palz_lbvh_normal_bounds": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_solve_primary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_image_fixed_to_float": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "palz_spectral_resolve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_photon_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_tuple_table_create": (C.c_int, [c_void_pp]),
    "palz_tuple_table_destroy": (C.c_int, [C.c_void_p]),
    "palz_tuple_table_build": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_void_p,
                                         C.c_void_p]),
    "palz_tuple_table_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_uint32)] * 4),
    "palz_lookup_flatten": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_solve_specular": (C.c_int, [C.c_int, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    "palz_renderer_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "palz_renderer_destroy": (C.c_int, [C.c_void_p]),
    "palz_renderer_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_renderer_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_renderer_set_camera": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_renderer_load_scene": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "palz_renderer_load_obj": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p]),
    "palz_renderer_step": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_renderer_clear": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_renderer_resolve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_renderer_buffers": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_mesh_load_obj": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p]),
    "palz_mesh_triangle_count": (C.c_int, [C.c_void_p, C.c_void_p]),
    "palz_mesh_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_mesh_destroy": (C.c_int, [C.c_void_p]),
    "palz_path_cuts_create": (C.c_int, [C.c_void_p]),
    "palz_path_cuts_destroy": (C.c_int, [C.c_void_p]),
    "palz_path_cuts_enumerate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                           C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_interval_cull": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "palz_affine_cull": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "palz_adjoint_irradiance": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    "palz_batched_create": (C.c_int, [c_void_pp]),
    "palz_batched_destroy": (C.c_int, [C.c_void_p]),
    "palz_batched_prepare": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "palz_batched_order": (C.c_int, [C.c_void_p, c_void_pp, c_void_pp, c_void_pp]),
    "palz_batched_set_config": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "palz_batched_get_config": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "palz_batched_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "palz_batched_solve_graph": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "palz_batched_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "palz_batched_path_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "palz_batched_set_branches": (C.c_int, [C.c_void_p, C.c_int]),
    "palz_intersect_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_float,
                                      C.c_void_p, C.c_void_p]),
    "palz_batched_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "palz_batched_timings": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    "palz_batched_set_culling": (C.c_int, [C.c_void_p, C.c_int]),
    "palz_batched_set_fused": (C.c_int, [C.c_void_p, C.c_int]),
    "palz_batched_get_fused": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "palz_batched_set_tune_cache": (C.c_int, [C.c_void_p, C.c_char_p]),
    "palz_batched_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int,
                                        C.POINTER(C.c_int32), C.c_void_p]),
    "palz_lbvh_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "palz_lbvh_timings": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int,
                                    C.POINTER(C.c_int)]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library (loads it on first use)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m palzilla_amd.build` "
                "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def symbols() -> list[str]:
    return list(_SIGS)


def call(name: str, *args) -> None:
    """Call an int-status entry point and raise PalzError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().palz_last_error().decode(errors="replace")
        raise PalzError(rc, name, msg)
