"""Pins of the CPU oracle's LBVH path (CPU only).

Each check is independent of the oracle's own formula: the reference's naive
Morton encoder (minimum_lbvh.h:149-160) against the bit-twiddling one, numpy's
stable argsort against the oracle's sort, a top-down split-at-first-max-delta
builder against the bottom-up Apetrei climb, and brute-force boxes, leaf
coverage, parent links and escape links against the built tree.
"""
import os

import numpy as np
import pytest

import oracle
from palzilla_amd import scenes
from palzilla_amd import types as T


def test_morton_splat_matches_naive_encoder():
    rng = np.random.default_rng(0)
    for x in range(8):
        for y in range(8):
            for z in range(8):
                assert oracle.morton_splat(x, y, z) == oracle.morton_naive(x, y, z)
    for x, y, z in rng.integers(0, 1 << 21, size=(2000, 3)):
        assert oracle.morton_splat(int(x), int(y), int(z)) == oracle.morton_naive(int(x), int(y), int(z))


def _decode(code: int):
    """Inverse of the 3-way interleave by plain bit loops."""
    xyz = [0, 0, 0]
    for b in range(63):
        if code >> b & 1:
            xyz[b % 3] |= 1 << (b // 3)
    return tuple(xyz)


def test_morton_keys_known_answers():
    # A single triangle per centroid; the scene box is fixed to [0,1]^3 by two
    # corner triangles so the quantisation is exact in float32.
    box = np.zeros(1, dtype=T.AABB)[0]
    box["lower"] = (0, 0, 0)
    box["upper"] = (1, 1, 1)
    pts = np.array([[0.5, 0.25, 0.75], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.125, 0.5, 0.0625]],
                   dtype=np.float32)
    tris = np.repeat(pts, 3, axis=0).reshape(-1, 9)  # degenerate triangles at the points
    k32 = oracle.morton_keys(tris, box, bits=32)
    k30 = oracle.morton_keys(tris, box, bits=30)
    for p, a, b in zip(pts, k32, k30):
        q21 = tuple(min(int(v * 2 ** 21), 2 ** 21 - 1) for v in p)
        q10 = tuple(min(int(v * 2 ** 10), 2 ** 10 - 1) for v in p)
        assert int(a) == oracle.morton_naive(*q21) >> 31
        assert _decode(int(b)) == q10


def test_morton_keys_z_order_monotone():
    xs = np.linspace(0.0, 1.0, 200, dtype=np.float32)
    pts = np.stack([xs, np.full_like(xs, 0.3), np.full_like(xs, 0.7)], axis=1)
    tris = np.repeat(pts, 3, axis=0).reshape(-1, 9)
    box = np.zeros(1, dtype=T.AABB)[0]
    box["lower"] = (0, 0, 0)
    box["upper"] = (1, 1, 1)
    for bits in (30, 32):
        k = oracle.morton_keys(tris, box, bits=bits)
        assert np.all(np.diff(k.astype(np.int64)) >= 0)


@pytest.mark.parametrize("start,end", [(0, 32), (0, 64), (8, 40), (3, 29)])
def test_stable_sort_matches_numpy(start, end):
    rng = np.random.default_rng(start * 100 + end)
    w = rng.integers(0, 2 ** 63, size=5000, dtype=np.uint64)
    w[::7] = w[3]  # duplicates
    v = np.arange(w.size, dtype=np.uint32)
    mask = np.uint64((1 << (end - start)) - 1) if end - start < 64 else np.uint64(2 ** 64 - 1)
    key = (w >> np.uint64(start)) & mask
    order = np.argsort(key, kind="stable")
    ws, vs = oracle.stable_sort_u64(w, start, end, values=v)
    assert np.array_equal(ws, w[order])
    assert np.array_equal(vs, v[order])


def _check_tree(tris, nodes, root, words, deltas):
    n = tris.shape[0]
    tris = tris.reshape(-1, 3, 3)
    assert oracle.validate_lbvh(nodes, root, deltas) if n > 1 else True
    if n == 1:
        assert root == (0 | T.NODE_LEAF_BIT)
        return
    seen = np.zeros(n, dtype=np.int32)
    parents = {root: T.NODE_INVALID}

    def box_of(node):
        if node & T.NODE_LEAF_BIT:
            t = tris[node & 0x7FFFFFFF]
            seen[node & 0x7FFFFFFF] += 1
            return t.min(axis=0), t.max(axis=0), [node & 0x7FFFFFFF]
        nd = nodes[node]
        lo, hi, leaves = None, None, []
        for c in range(2):
            ch = int(nd["children"][c])
            clo, chi, cl = box_of(ch)
            assert np.array_equal(nd["aabbs"][c]["lower"], clo)
            assert np.array_equal(nd["aabbs"][c]["upper"], chi)
            if not ch & T.NODE_LEAF_BIT:
                assert int(nodes[ch]["parent"]) == node
            lo = clo if lo is None else np.minimum(lo, clo)
            hi = chi if hi is None else np.maximum(hi, chi)
            leaves += cl
        return lo, hi, leaves

    import sys
    sys.setrecursionlimit(10000)
    box_of(root)
    assert np.all(seen == 1), "every triangle is exactly one leaf"
    assert int(nodes[root]["parent"]) == T.NODE_INVALID
    # escape link = upper end of the node's sorted-leaf range (or invalid at the right spine)
    pos = {int(w >> np.uint64(32)): i for i, w in enumerate(words)}

    def rng_of(node):
        if node & T.NODE_LEAF_BIT:
            p = pos[node & 0x7FFFFFFF]
            return p, p
        a, _ = rng_of(int(nodes[node]["children"][0]))
        _, b = rng_of(int(nodes[node]["children"][1]))
        ctx = int(nodes[node]["context"])
        assert ctx == (b if b < n - 1 else 0xFFFFFFFF)
        assert a <= node < b, "internal node i splits leaves i and i+1"
        return a, b

    assert rng_of(root) == (0, n - 1)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 257, 1000])
@pytest.mark.parametrize("bits", [30, 32])
def test_bottom_up_equals_top_down(n, bits):
    tris = scenes.triangle_soup(n, seed=n + bits)
    if n > 8:  # duplicate centroids exercise delta == 0 ties
        tris[n // 2: n // 2 + 4] = tris[n // 3]
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=bits)
    if n > 1:
        td_nodes, td_root = oracle.lbvh_topdown(tris, words, deltas)
        assert td_root == root
        assert td_nodes.tobytes() == nodes.tobytes()
    _check_tree(tris, nodes, root, words, deltas)
    assert np.array_equal(box["lower"], tris.reshape(-1, 3).min(axis=0))
    assert np.array_equal(box["upper"], tris.reshape(-1, 3).max(axis=0))


def test_sorted_words_are_key_index_order():
    tris = scenes.triangle_soup(3000, seed=11)
    tris[100:140] = tris[7]
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=30)
    keys = (words & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    idx = (words >> np.uint64(32)).astype(np.uint64)
    composite = keys << np.uint64(32) | idx
    assert np.all(composite[1:] > composite[:-1])
    assert sorted(idx.tolist()) == list(range(3000))
    # deltas are the highest differing bit position + 1 of neighbouring keys
    k = keys.astype(np.int64)
    x = k[:-1] ^ k[1:]
    expect = np.array([int(v).bit_length() for v in x], dtype=np.uint8)
    assert np.array_equal(deltas, expect)


@pytest.mark.parametrize("bits", [30, 32])
def test_reference_mesh_fixture(bits):
    """The reference's own mesh (bin/assets/table/table.obj): the bottom-up
    build equals the independent top-down builder and passes every tree
    invariant; exact-duplicate centroids of the cylinder caps exercise ties."""
    tris = scenes.load_obj(os.path.join(os.path.dirname(__file__), "golden", "table.obj"))
    assert tris.shape == (3068, 9)
    nodes, root, words, deltas, box = oracle.lbvh(tris, bits=bits)
    td_nodes, td_root = oracle.lbvh_topdown(tris, words, deltas)
    assert td_root == root and td_nodes.tobytes() == nodes.tobytes()
    _check_tree(tris, nodes, root, words, deltas)


def _subtree_leaves(nodes, node):
    out, stack = [], [node]
    while stack:
        x = int(stack.pop())
        if x & 0x80000000:
            out.append(x & 0x7FFFFFFF)
        else:
            stack.extend(nodes["children"][x])
    return out


def random_attribs(n, seed):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=T.TRIANGLE_ATTRIB)
    a["material"] = 2
    a["shading_normals"] = rng.normal(size=(n, 3, 3)).astype(np.float32)
    return a


@pytest.mark.parametrize("n", [2, 3, 17, 1000])
def test_normal_bounds_brute_force(n):
    # every slot holds the box of the shading normals under that child: the
    # serial restatement of main.cpp:248-295 against a direct subtree walk
    tris = scenes.triangle_soup(n, seed=n + 11)
    nodes, root, *_ = oracle.lbvh(tris, option=T.BUILD_OPTION_USE_NORMAL, bits=32)
    attribs = random_attribs(n, n)
    nb = oracle.normal_bounds(nodes, attribs)
    assert np.all(nb["counter"] == 2)
    for i in range(n - 1):
        for c in range(2):
            sn = attribs["shading_normals"][_subtree_leaves(nodes, nodes["children"][i][c])].reshape(-1, 3)
            box = nb["normal_bounds"][i][c]
            assert np.array_equal(box["lower"], sn.min(axis=0)), (i, c)
            assert np.array_equal(box["upper"], sn.max(axis=0)), (i, c)


@pytest.mark.parametrize("n,threads", [(2, 4), (257, 3), (4096, 8), (100_003, 8), (300_000, 16)])
def test_parallel_cpu_build_equals_the_serial_one(n, threads):
    # the threaded CPU baseline (radix sort + atomic climb) builds the same tree
    tris = scenes.triangle_soup(n, seed=n)
    if n > 1000:
        tris[n // 3: n // 3 + 500] = tris[n // 3]  # a long run of equal keys
    for bits in (30, 32):
        want, root, *_ = oracle.lbvh(tris, bits=bits)
        got, groot = oracle.lbvh_parallel(tris, bits=bits, threads=threads)
        assert groot == root
        assert got.tobytes() == want.tobytes()
