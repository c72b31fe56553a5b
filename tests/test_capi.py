"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls)."""
import ctypes
import glob
import os
import re

import palzilla_amd
from palzilla_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(palz_[a-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(palzilla_amd.LIB_PATH)
    declared = _declared()
    assert len(declared) > 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    assert _declared() == set(_lib.symbols())


def test_error_reporting_without_gpu_work():
    lib = _lib.lib()
    assert lib.palz_version() >= 100
    # invalid arguments are rejected before any HIP call
    rc = lib.palz_sort_u64(None, None, None, None, None, 0, 0, 32, None)
    assert rc == 1
    assert b"sorter" in lib.palz_last_error()


def test_new_entry_points_reject_bad_arguments_before_gpu_work():
    # path cuts, renderer and mesh entry points check their arguments first
    lib = _lib.lib()
    INVALID = 1
    assert lib.palz_path_cuts_create(None) == INVALID
    pc = ctypes.c_void_p()
    assert lib.palz_path_cuts_create(ctypes.byref(pc)) == 0
    try:
        desc = ctypes.create_string_buffer(104)
        count = ctypes.c_void_p(8)  # never dereferenced: rejected before any launch
        for event in (0, 4 << 24 | 0xF):  # chain length 0 and 4
            assert lib.palz_path_cuts_enumerate(pc, desc, None, event, ctypes.c_float(1.5), ctypes.c_float(1.5), 0,
                                                None, 0, count, None, None) == INVALID
            assert b"chain length" in lib.palz_last_error()
        tt = 2 << 24 | 3
        assert lib.palz_path_cuts_enumerate(pc, desc, None, tt, ctypes.c_float(1.6), ctypes.c_float(1.5), 0,
                                            None, 0, count, None, None) == INVALID  # eta_lo > eta_hi
        assert lib.palz_path_cuts_enumerate(pc, desc, None, tt, ctypes.c_float(1.5), ctypes.c_float(1.6), 0,
                                            None, 16, count, None, None) == INVALID  # capacity without d_paths
        assert lib.palz_path_cuts_enumerate(pc, None, None, tt, ctypes.c_float(1.5), ctypes.c_float(1.6), 0,
                                            None, 0, count, None, None) == INVALID
    finally:
        lib.palz_path_cuts_destroy(pc)
    lam = (ctypes.c_float * 2)(400.0, 600.0)
    r = ctypes.c_void_p()
    assert lib.palz_renderer_create(0, 16, lam, lam, 2, ctypes.byref(r)) == INVALID
    assert lib.palz_renderer_create(16, 16, lam, lam, 0, ctypes.byref(r)) == INVALID
    assert lib.palz_renderer_create(16, 16, None, lam, 2, ctypes.byref(r)) == INVALID
    assert lib.palz_renderer_create(1 << 13, 1 << 13, lam, lam, 2, ctypes.byref(r)) == INVALID
    assert lib.palz_renderer_step(None, None) == INVALID
    assert lib.palz_renderer_resolve(None, None, None) == INVALID
    m = ctypes.c_void_p()
    assert lib.palz_mesh_load_obj(b"/nonexistent/x.obj", 0, ctypes.byref(m)) == INVALID
    assert lib.palz_mesh_load_obj(b"/nonexistent/x.obj", 7, ctypes.byref(m)) == INVALID
    assert b"default material" in lib.palz_last_error()
