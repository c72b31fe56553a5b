"""ctypes binding of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / CPU baseline.  The product library
never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_SIGS = {
    "orc_scene_aabb": (None, [_P, C.c_int, _P]),
    "orc_morton_keys": (None, [_P, C.c_int, _P, C.c_uint32, C.c_int, _P]),
    "orc_morton_naive": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_morton_splat": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_stable_sort_u64": (None, [_P, _P, C.c_size_t, C.c_int, C.c_int]),
    "orc_deltas": (None, [_P, C.c_int, _P]),
    "orc_build_lbvh": (None, [_P, C.c_int, _P, _P, _P, _P]),
    "orc_lbvh": (None, [_P, C.c_int, C.c_uint32, C.c_int, _P, _P, _P, _P, _P]),
    "orc_lbvh_topdown": (None, [_P, C.c_int, _P, _P, _P, _P]),
    "orc_validate_lbvh": (C.c_int, [_P, C.c_uint32, _P]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `python oracle/build.py`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def scene_aabb(tris: np.ndarray) -> np.ndarray:
    from palzilla_amd import types as T
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    out = np.zeros(1, dtype=T.AABB)
    lib().orc_scene_aabb(_ptr(tris), tris.shape[0], _ptr(out))
    return out[0]


def morton_keys(tris: np.ndarray, aabb: np.ndarray, option: int = 0, bits: int = 32) -> np.ndarray:
    from palzilla_amd import types as T
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    box = np.array([aabb], dtype=T.AABB)
    keys = np.zeros(tris.shape[0], dtype=np.uint32)
    lib().orc_morton_keys(_ptr(tris), tris.shape[0], _ptr(box), option, bits, _ptr(keys))
    return keys


def stable_sort_u64(words: np.ndarray, start_bit: int, end_bit: int, values=None):
    w = np.ascontiguousarray(words, dtype=np.uint64).copy()
    v = None if values is None else np.ascontiguousarray(values, dtype=np.uint32).copy()
    lib().orc_stable_sort_u64(_ptr(w), _ptr(v), w.size, start_bit, end_bit)
    return w if v is None else (w, v)


def lbvh(tris: np.ndarray, option: int = 0, bits: int = 32):
    """Full CPU build: returns (nodes, root, sorted_words, deltas, scene_aabb)."""
    from palzilla_amd import types as T
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    n = tris.shape[0]
    nodes = np.zeros(max(n - 1, 1), dtype=T.INTERNAL_NODE)
    root = np.zeros(1, dtype=np.uint32)
    words = np.zeros(n, dtype=np.uint64)
    deltas = np.zeros(max(n - 1, 1), dtype=np.uint8)
    box = np.zeros(1, dtype=T.AABB)
    lib().orc_lbvh(_ptr(tris), n, option, bits, _ptr(nodes), _ptr(root), _ptr(words), _ptr(deltas),
                   _ptr(box))
    return nodes[: n - 1], int(root[0]), words, deltas[: n - 1], box[0]


def lbvh_topdown(tris: np.ndarray, sorted_words: np.ndarray, deltas: np.ndarray):
    from palzilla_amd import types as T
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    n = tris.shape[0]
    nodes = np.zeros(max(n - 1, 1), dtype=T.INTERNAL_NODE)
    root = np.zeros(1, dtype=np.uint32)
    w = np.ascontiguousarray(sorted_words, dtype=np.uint64)
    d = np.ascontiguousarray(deltas, dtype=np.uint8)
    lib().orc_lbvh_topdown(_ptr(tris), n, _ptr(w), _ptr(d), _ptr(nodes), _ptr(root))
    return nodes[: n - 1], int(root[0])


def build_lbvh(tris: np.ndarray, sorted_words: np.ndarray, deltas: np.ndarray):
    from palzilla_amd import types as T
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    n = tris.shape[0]
    nodes = np.zeros(max(n - 1, 1), dtype=T.INTERNAL_NODE)
    root = np.zeros(1, dtype=np.uint32)
    lib().orc_build_lbvh(_ptr(tris), n, _ptr(np.ascontiguousarray(sorted_words, dtype=np.uint64)),
                         _ptr(np.ascontiguousarray(deltas, dtype=np.uint8)), _ptr(nodes), _ptr(root))
    return nodes[: n - 1], int(root[0])


def validate_lbvh(nodes: np.ndarray, root: int, deltas: np.ndarray) -> bool:
    return bool(lib().orc_validate_lbvh(_ptr(np.ascontiguousarray(nodes)), root,
                                        _ptr(np.ascontiguousarray(deltas, dtype=np.uint8))))


def morton_naive(x: int, y: int, z: int) -> int:
    return int(lib().orc_morton_naive(x, y, z))


def morton_splat(x: int, y: int, z: int) -> int:
    return int(lib().orc_morton_splat(x, y, z))
