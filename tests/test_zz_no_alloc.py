"""palz_sort_u64_noalloc (the sort with its scratch in a file-scope device
array) through the C ABI: equal, bit for bit, to palz_sort_u64 (a sorter's
own scratch) and to numpy's stable argsort, on two calls in a row, the second
with a smaller input, and a size above the array's maximum rejected."""
import ctypes as C

import numpy as np
import pytest

from palzilla_amd._lib import call, lib
from palzilla_amd.device import DeviceBuffer
from palzilla_amd.lbvh import OnesweepSort

pytestmark = pytest.mark.gpu

MAX_N = 1 << 24
PALZ_ERR_CAPACITY = 4


def _noalloc(words, start, end, values=None):
    n = words.size
    dw, dt = DeviceBuffer.from_array(words), DeviceBuffer(8 * n)
    dv = dtv = None
    if values is not None:
        dv, dtv = DeviceBuffer.from_array(values), DeviceBuffer(4 * n)
    call("palz_sort_u64_noalloc", C.c_void_p(dw.ptr), C.c_void_p(dt.ptr), C.c_void_p(dv.ptr if dv else None),
         C.c_void_p(dtv.ptr if dtv else None), C.c_uint32(n), int(start), int(end), C.c_void_p(None))
    return dw.download(np.uint64, n), (dv.download(np.uint32, n) if dv else None)


def _sorter(sorter, words, start, end, values=None):
    n = words.size
    dw, dt = DeviceBuffer.from_array(words), DeviceBuffer(8 * n)
    dv = dtv = None
    if values is not None:
        dv, dtv = DeviceBuffer.from_array(values), DeviceBuffer(4 * n)
    sorter.sort(dw.ptr, dt.ptr, n, start, end, values=dv.ptr if dv else 0, tmp_values=dtv.ptr if dtv else 0)
    return dw.download(np.uint64, n), (dv.download(np.uint32, n) if dv else None)


@pytest.mark.parametrize("calls", [[(3_000_017, 0, 40, True), (200_003, 8, 32, False)],
                                   [(MAX_N, 0, 32, False), (65_541, 0, 24, True)]])
def test_noalloc_sort_equals_the_sorter_on_two_calls(calls):
    sorter = OnesweepSort()
    rng = np.random.default_rng(2024)
    for n, start, end, with_values in calls:
        print(f"noalloc sort: n={n} bits=[{start},{end}) values={with_values}", flush=True)
        w = rng.integers(0, 2 ** 64 - 1, size=n, dtype=np.uint64)
        w[::3] = w[7]  # runs of equal keys: stability
        v = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32) if with_values else None
        got_w, got_v = _noalloc(w, start, end, v)
        ref_w, ref_v = _sorter(sorter, w, start, end, v)
        assert np.array_equal(got_w, ref_w)
        order = np.argsort((w >> np.uint64(start)) & np.uint64((1 << (end - start)) - 1), kind="stable")
        assert np.array_equal(got_w, w[order])
        if with_values:
            assert np.array_equal(got_v, ref_v) and np.array_equal(got_v, v[order])


def test_noalloc_sort_rejects_more_than_its_maximum():
    d = DeviceBuffer(64)
    rc = lib().palz_sort_u64_noalloc(C.c_void_p(d.ptr), C.c_void_p(d.ptr), None, None, C.c_uint32(MAX_N + 1), 0,
                                     32, None)
    assert rc == PALZ_ERR_CAPACITY
    # the rejected call touched nothing: a following call still sorts correctly
    w = np.arange(1000, dtype=np.uint64)[::-1].copy()
    got, _ = _noalloc(w, 0, 16)
    assert np.array_equal(got, np.sort(w))
