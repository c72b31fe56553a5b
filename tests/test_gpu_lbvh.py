"""GPU parity of the onesweep sort and the LBVH build (bit-exact vs the CPU
oracle and numpy's stable argsort)."""
import os

import numpy as np
import pytest

import oracle
from palzilla_amd import scenes
from palzilla_amd import types as T
from palzilla_amd.device import DeviceBuffer
from palzilla_amd.lbvh import BVHGPUBuilder, OnesweepSort

pytestmark = pytest.mark.gpu


def _gpu_sort(sorter, words, start, end, values=None):
    n = words.size
    dw = DeviceBuffer.from_array(words)
    dt = DeviceBuffer(8 * max(n, 1))
    dv = dtv = None
    if values is not None:
        dv = DeviceBuffer.from_array(values)
        dtv = DeviceBuffer(4 * max(n, 1))
    sorter.sort(dw.ptr, dt.ptr, n, start, end, values=dv.ptr if dv else 0,
                tmp_values=dtv.ptr if dtv else 0)
    out = dw.download(np.uint64, n)
    if values is None:
        return out
    return out, dv.download(np.uint32, n)


@pytest.mark.parametrize("n", [1, 63, 4095, 4096, 4097, 65539, 1 << 20])
@pytest.mark.parametrize("start,end", [(0, 32), (0, 64), (3, 29)])
def test_sort_matches_stable_argsort(sorter, n, start, end):
    print(f"sort n={n} bits=[{start},{end})", flush=True)
    rng = np.random.default_rng(n + start * 7 + end)
    w = rng.integers(0, 2 ** 64 - 1, size=n, dtype=np.uint64)
    if n > 10:
        w[:: 5] = w[1]  # long runs of equal keys test stability
    v = np.arange(n, dtype=np.uint32)[::-1].copy()
    width = end - start
    mask = np.uint64(2 ** 64 - 1) if width == 64 else np.uint64((1 << width) - 1)
    order = np.argsort((w >> np.uint64(start)) & mask, kind="stable")
    got_w, got_v = _gpu_sort(sorter, w, start, end, values=v)
    assert np.array_equal(got_w, w[order])
    assert np.array_equal(got_v, v[order])
    got_only = _gpu_sort(sorter, w, start, end)
    assert np.array_equal(got_only, w[order])


@pytest.mark.parametrize("n", [1, 4097, 65539, 1 << 20])
@pytest.mark.parametrize("start,end", [(0, 32), (0, 30), (5, 21)])
def test_sort_u32_matches_stable_argsort(sorter, n, start, end):
    # the KeyValueSoA32 overload (tinyhiponesweep.h:592): u32 keys, optional u32 values
    print(f"sort32 n={n} bits=[{start},{end})", flush=True)
    rng = np.random.default_rng(3 * n + start + end)
    k = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        k[::7] = k[2]
    v = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    order = np.argsort((k >> np.uint32(start)) & np.uint32((1 << (end - start)) - 1), kind="stable")
    dk, dt = DeviceBuffer.from_array(k), DeviceBuffer(4 * n)
    dv, dtv = DeviceBuffer.from_array(v), DeviceBuffer(4 * n)
    sorter.sort32(dk.ptr, dt.ptr, n, start, end, values=dv.ptr, tmp_values=dtv.ptr)
    assert np.array_equal(dk.download(np.uint32, n), k[order])
    assert np.array_equal(dv.download(np.uint32, n), v[order])
    dk2 = DeviceBuffer.from_array(k)
    sorter.sort32(dk2.ptr, dt.ptr, n, start, end)
    assert np.array_equal(dk2.download(np.uint32, n), k[order])


def test_sort_skewed_digits(sorter):
    # all keys share the high digits: one bin holds the whole tile in late passes
    n = 300_000
    rng = np.random.default_rng(5)
    w = rng.integers(0, 256, size=n, dtype=np.uint64) | np.uint64(0xABCD0000)
    order = np.argsort(w & np.uint64(0xFFFFFFFF), kind="stable")
    assert np.array_equal(_gpu_sort(sorter, w, 0, 32), w[order])


def test_sort_repeat_calls_reuse_lookback(sorter):
    rng = np.random.default_rng(9)
    for i in range(6):
        n = int(rng.integers(1, 200_000))
        w = rng.integers(0, 2 ** 40, size=n, dtype=np.uint64)
        order = np.argsort(w & np.uint64(0xFFFFFFFFFF), kind="stable")
        assert np.array_equal(_gpu_sort(sorter, w, 0, 40), w[order])


def _gpu_lbvh(sorter, tris, bits, option=0):
    dt = DeviceBuffer.from_array(tris)
    b = BVHGPUBuilder()
    b.build(dt.ptr, tris.shape[0], option, sorter, morton_bits=bits)
    return b


@pytest.mark.parametrize("n", [1, 2, 3, 17, 256, 257, 4097, 100_003, 1 << 20])
@pytest.mark.parametrize("bits", [30, 32])
def test_lbvh_bit_exact(sorter, n, bits):
    print(f"lbvh n={n} bits={bits}", flush=True)
    tris = scenes.triangle_soup(n, seed=n * 3 + bits)
    if n > 8:
        tris[n // 2: n // 2 + 5] = tris[n // 3]
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=bits)
    b = _gpu_lbvh(sorter, tris, bits)
    g_box = b.download_scene_aabb()
    assert g_box.tobytes() == box.tobytes()
    g_words, g_deltas = b.download_sorted()
    assert np.array_equal(g_words, words)
    assert np.array_equal(g_deltas, deltas)
    assert b.download_root() == root
    g_nodes = b.download_nodes()
    assert g_nodes.tobytes() == nodes.tobytes()


@pytest.mark.parametrize("n,run", [(5000, 5000), (100_003, 20_000), (300_000, 3000), (40_000, 600)])
def test_lbvh_long_equal_key_runs(sorter, n, run):
    # a run of identical triangles gives equal Morton keys, whose splits all
    # have delta 0: a right-leaning chain that no chunk of leaves can close
    print(f"lbvh equal-key run n={n} run={run}", flush=True)
    tris = scenes.triangle_soup(n, seed=n + run)
    start = (n - run) // 3
    tris[start:start + run] = tris[start]
    for bits in (30, 32):
        nodes, root, words, deltas, box = oracle.lbvh(tris, bits=bits)
        assert np.count_nonzero(deltas == 0) >= run - 1
        b = _gpu_lbvh(sorter, tris, bits)
        g_words, g_deltas = b.download_sorted()
        assert np.array_equal(g_deltas, deltas)
        assert b.download_root() == root
        assert b.download_nodes().tobytes() == nodes.tobytes()


def test_lbvh_use_normal_option(sorter):
    tris = scenes.triangle_soup(5000, seed=77)
    nodes, root, words, deltas, box = oracle.lbvh(tris, option=T.BUILD_OPTION_USE_NORMAL, bits=32)
    b = _gpu_lbvh(sorter, tris, 32, option=T.BUILD_OPTION_USE_NORMAL)
    g_words, _ = b.download_sorted()
    assert np.array_equal(g_words, words)
    assert b.download_nodes().tobytes() == nodes.tobytes()


def test_lbvh_rebuild_and_unaligned_input(sorter):
    # a second build of a different size through the same builder, from a
    # triangle array that is not 16-byte aligned (staging falls back to dwords)
    b = BVHGPUBuilder()
    for n, off in ((5000, 0), (777, 1)):
        tris = scenes.triangle_soup(n, seed=n)
        buf = np.zeros((n + 1) * 9, dtype=np.float32)
        buf[off * 9: off * 9 + n * 9] = tris.reshape(-1)
        dt = DeviceBuffer.from_array(buf)
        b.build(dt.ptr + off * 36, n, 0, sorter, morton_bits=32)
        nodes, root, *_ = oracle.lbvh(tris, bits=32)
        assert b.download_root() == root
        assert b.download_nodes().tobytes() == nodes.tobytes()


def test_lbvh_heightfield_grid(sorter):
    # structured input: a regular triangulated grid (many equal-delta runs)
    m = 64
    x, z = np.meshgrid(np.linspace(-1, 1, m + 1), np.linspace(-1, 1, m + 1), indexing="ij")
    y = 0.5 + 0.02 * np.sin(3 * x) * np.cos(5 * z)
    P = np.stack([x, y, z], axis=-1)
    tris = []
    for i in range(m):
        for j in range(m):
            a, b_, c, d = P[i, j], P[i + 1, j], P[i + 1, j + 1], P[i, j + 1]
            tris += [np.concatenate([a, b_, c]), np.concatenate([a, c, d])]
    tris = np.array(tris, dtype=np.float32)
    for bits in (30, 32):
        nodes, root, *_ = oracle.lbvh(tris, bits=bits)
        b = _gpu_lbvh(sorter, tris, bits)
        assert b.download_root() == root
        assert b.download_nodes().tobytes() == nodes.tobytes()


@pytest.mark.parametrize("bits", [30, 32])
def test_lbvh_reference_mesh_bit_exact(sorter, bits):
    """The reference's own mesh (bin/assets/table/table.obj), bit for bit."""
    tris = scenes.load_obj(os.path.join(os.path.dirname(__file__), "golden", "table.obj"))
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=bits)
    b = _gpu_lbvh(sorter, tris, bits)
    assert b.download_scene_aabb().tobytes() == box.tobytes()
    g_words, g_deltas = b.download_sorted()
    assert np.array_equal(g_words, words) and np.array_equal(g_deltas, deltas)
    assert b.download_root() == root
    assert b.download_nodes().tobytes() == nodes.tobytes()


@pytest.mark.parametrize("n,option,run", [(2, 0, 0), (3, 0, 0), (257, 0, 0), (4097, 1, 0), (100_003, 1, 0),
                                          (40_000, 0, 600)])
def test_normal_bounds_bit_exact(sorter, n, option, run):
    # InternalNormalBound (main.cpp:248-295) on the GPU tree, against the
    # oracle's serial pass; `run` identical triangles exercise the
    # device-wide part of the build
    print(f"normal bounds n={n} option={option} run={run}", flush=True)
    tris = scenes.triangle_soup(n, seed=n + 5)
    if run:
        tris[n // 3:n // 3 + run] = tris[n // 3]
    b = BVHGPUBuilder()
    dt = DeviceBuffer.from_array(tris)
    b.build(dt.ptr, n, option, sorter, morton_bits=32)
    rng = np.random.default_rng(n)
    attribs = np.zeros(n, dtype=T.TRIANGLE_ATTRIB)
    attribs["material"] = 2
    attribs["shading_normals"] = rng.normal(size=(n, 3, 3)).astype(np.float32)
    d_attr = DeviceBuffer.from_array(attribs)
    got = b.normal_bounds(d_attr.ptr).download(T.NORMAL_BOUND, n - 1)
    want = oracle.normal_bounds(b.download_nodes(), attribs)
    assert got.tobytes() == want.tobytes()
    assert np.all(got["counter"] == 2)


@pytest.fixture(scope="module")
def benchmark_soup():
    # the LBVH benchmark's 16M-triangle soup; its 4M build is the first 4M
    return scenes.triangle_soup(1 << 24, seed=2024)


@pytest.mark.parametrize("n", [1 << 22, 1 << 24])
def test_lbvh_bit_exact_benchmark_sizes(sorter, benchmark_soup, n):
    """The benchmark's own 4M and 16M builds (30-bit keys), bit for bit against
    the oracle: scene box, sorted words, deltas, root and every node."""
    print(f"lbvh benchmark size n={n}", flush=True)
    tris = np.ascontiguousarray(benchmark_soup[:n])
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=30)
    b = _gpu_lbvh(sorter, tris, 30)
    assert b.download_scene_aabb().tobytes() == box.tobytes()
    g_words, g_deltas = b.download_sorted()
    assert np.array_equal(g_words, words)
    assert np.array_equal(g_deltas, deltas)
    assert b.download_root() == root
    assert b.download_nodes().tobytes() == nodes.tobytes()
