"""The oracle's node-level path cuts (traverseAdmissibleNodes, pk.h:1760-2212),
pinned without its traversal: the set it emits per receiver point equals a
brute force over every K-tuple of triangles (leaf boxes, the closed-mesh side
tests restated in numpy float32), each tuple once; its order for K = 1 is
the tree's in-order leaf sequence reversed (a LIFO walk that pushes child 0
first); and no tuple whose chain the FP32 solver makes contribute is cut."""
import itertools

import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes

LEAF = 0x80000000


def _cmin(a):
    lo = a.min(axis=0)
    neg0 = ((a == 0) & np.signbit(a)).any(axis=0)
    return np.where((lo == 0) & neg0, np.float32(-0.0), lo).astype(np.float32)


def _cmax(a):
    hi = a.max(axis=0)
    pos0 = ((a == 0) & ~np.signbit(a)).any(axis=0)
    return np.where((hi == 0) & pos0, np.float32(0.0), hi).astype(np.float32)


def _hull(p3):
    return np.concatenate([_cmin(p3), _cmax(p3)])


def _front(tri, sn0, p):
    """pk.h:1850-1875 in float32, one operation at a time."""
    f = np.float32
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    ng = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]], f)
    d = ng[0] * sn0[0] + ng[1] * sn0[1] + ng[2] * sn0[2]
    ng = ng * d
    return not ((ng[0] * p[0] + ng[1] * p[1] + ng[2] * p[2]) < (ng[0] * tri[0][0] + ng[1] * tri[0][1] + ng[2] * tri[0][2]))


class Case:
    def __init__(self, scene, events):
        self.s = scene
        self.tris = scene.tris.reshape(-1, 3, 3).astype(np.float32)
        self.sn = scene.attribs["shading_normals"].astype(np.float32)
        self.nodes, self.root, _, _, _ = oracle.lbvh(scene.tris, bits=32)
        self.nb = oracle.normal_bounds(self.nodes, scene.attribs)
        lam = scene.lambdas.astype(np.float32)
        eta = np.float32(scene.cauchy_a) + np.float32(scene.cauchy_b) / (lam * lam)
        self.eta = (float(eta.min()), float(eta.max()))
        self.events = events
        self.vbox = np.stack([_hull(t) for t in self.tris])
        self.nbox = np.stack([_hull(n) for n in self.sn])

    def cuts(self, event, points):
        return oracle.path_cuts(self.s.tris, self.s.attribs, self.nodes, self.root, self.nb, points,
                                self.s.p_light, event, *self.eta)

    def brute(self, event, p_end):
        K = event >> 24
        pl = np.asarray(self.s.p_light, dtype=np.float32)
        n = self.tris.shape[0]
        first_ok = [_front(self.tris[t], self.sn[t][0], pl) for t in range(n)]
        last_ok = [_front(self.tris[t], self.sn[t][0], p_end) for t in range(n)]
        out = set()
        for tup in itertools.product(range(n), repeat=K):
            if any(tup[k] == tup[k + 1] for k in range(K - 1)):
                continue
            if not (first_ok[tup[0]] and last_ok[tup[-1]]):
                continue
            if oracle.tuple_boxes_ok(self.vbox[list(tup)], self.nbox[list(tup)], pl, p_end, event, *self.eta):
                out.add(tup)
        return out


CASES = {
    "slab_TT": (lambda: scenes.config2(cells=4, width=8, photons=1 << 10, n_lambdas=2), ["TT", "T"]),
    "mirror_R": (lambda: scenes.stack_scene(mirror_cells=3, photons=1 << 10, width=8, height=8, seed=4), ["R"]),
    "mirror_slab_3": (lambda: scenes.stack_scene(mirror_cells=2, layer="slab", layer_cells=2, ior=1.6, seed=6,
                                                 photons=1 << 10, width=8, height=8), ["RTT", "TRT"]),
}
_built = {}


def _case(name):
    if name not in _built:
        make, events = CASES[name]
        _built[name] = Case(make(), events)
    return _built[name]


@pytest.mark.parametrize("name", sorted(CASES))
def test_cut_set_equals_brute_force(name):
    c = _case(name)
    rng = np.random.default_rng(3)
    pts = c.s.points[rng.choice(c.s.points.shape[0], 5, replace=False)]
    for ev in c.events:
        event = CA.event_u32(ev)
        counts, tuples = c.cuts(event, pts)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        total_found = 0
        for q in range(pts.shape[0]):
            got = [tuple(int(x) for x in t) for t in tuples[off[q]:off[q + 1]]]
            assert len(got) == len(set(got)), f"{name} {ev}: a tuple emitted twice"
            want = c.brute(event, pts[q]["p"].astype(np.float32))
            assert set(got) == want, f"{name} {ev} point {q}: {len(set(got) ^ want)} tuples differ"
            total_found += len(got)
        assert total_found > 0, f"{name} {ev}: nothing admissible, the check is vacuous"


def _in_order_leaves(nodes, root):
    out = []

    def walk(n):
        if n & LEAF:
            out.append(n & ~LEAF)
            return
        walk(int(nodes[n]["children"][0]))
        walk(int(nodes[n]["children"][1]))

    walk(int(root))
    return out


def test_k1_order_is_reverse_in_order():
    c = _case("mirror_R")
    order = _in_order_leaves(c.nodes, c.root)
    rank = {t: i for i, t in enumerate(order)}
    counts, tuples = c.cuts(CA.event_u32("R"), c.s.points[:16])
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    assert counts.sum() > 0
    for q in range(16):
        got = tuples[off[q]:off[q + 1], 0].tolist()
        assert got == sorted(got, key=lambda t: -rank[t])


def test_points_without_a_hit_get_nothing():
    c = _case("slab_TT")
    pts = c.s.points[:4].copy()
    pts["R"][1] = 0.0
    counts, _ = c.cuts(CA.event_u32("TT"), pts)
    assert counts[1] == 0 and counts[[0, 2, 3]].sum() > 0


@pytest.mark.parametrize("name,ev", [("slab_TT", "TT"), ("mirror_R", "R"), ("mirror_slab_3", "RTT")])
def test_contributing_chains_survive_the_cuts(name, ev):
    # every tuple whose chain the FP32 solver (started at the centroid) makes
    # contribute at a point is among that point's cut tuples
    c = _case(name)
    event = CA.event_u32(ev)
    K = event >> 24
    n = c.tris.shape[0]
    pts_idx = np.arange(0, c.s.points.shape[0], 7)
    counts, tuples = c.cuts(event, c.s.points[pts_idx])
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    allt = np.array(list(itertools.product(range(n), repeat=K)), dtype=np.int32)
    paths = np.zeros(allt.shape[0] * pts_idx.size, dtype=CA.PATH)
    paths["pixel"] = np.repeat(pts_idx, allt.shape[0])
    paths["event"] = event
    paths["tris"][:] = -1
    paths["tris"][:, :K] = np.tile(allt, (pts_idx.size, 1))
    paths["params"][:, :2 * K] = np.float32(1.0) / np.float32(3.0)
    hs = oracle.HostScene(c.s, c.nodes, c.root)
    res = oracle.solve_paths(hs, paths, threads=8).reshape(paths.shape[0], -1)
    contributing = (res["status"] == CA.PATH_CONTRIBUTES).any(axis=1)
    assert contributing.sum() > 0, "no contributing chain: the check is vacuous"
    for j in np.nonzero(contributing)[0]:
        q = j // allt.shape[0]
        t = tuple(int(x) for x in paths["tris"][j, :K])
        cut = {tuple(int(x) for x in u) for u in tuples[off[q]:off[q + 1]]}
        assert t in cut, f"{name} {ev}: contributing tuple {t} at point {pts_idx[q]} was cut"
