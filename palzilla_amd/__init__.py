"""palzilla_amd — the caustic hot path of Palzilla, rebuilt for MI355X (gfx950).

The compute lives in libpalz.so (hand-written HIP kernels behind the C ABI of
include/palz.h); this package is the host-side mirror of the reference's
classes plus scene generators for the benchmark configurations.
"""
from ._lib import LIB_PATH, PalzError, lib, symbols  # noqa: F401

__all__ = ["LIB_PATH", "PalzError", "lib", "symbols"]
