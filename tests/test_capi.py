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
