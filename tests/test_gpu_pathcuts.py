"""GPU tests of the node-level path cuts (palz_path_cuts_enumerate): records
byte-identical to the oracle's traversal (tuples, order, every field), the
capacity and count-only modes, containment of every tuple the photon
pipeline makes contribute, and a batched solve of the enumerated paths
bit-exact against the oracle's solve of the same paths."""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

SMALL = {
    "slab": (lambda: scenes.config2(cells=4, width=8, photons=1 << 10, n_lambdas=2), ["TT", "T"]),
    "mirror": (lambda: scenes.stack_scene(mirror_cells=3, photons=1 << 10, width=8, height=8, seed=4), ["R", "RR"]),
    "mirror_slab": (lambda: scenes.stack_scene(mirror_cells=2, layer="slab", layer_cells=2, ior=1.6, seed=6,
                                               photons=1 << 10, width=4, height=4), ["RTT", "RT"]),
    "pool": (lambda: scenes.stack_scene(layer="pool", layer_cells=6, ior=1.33, seed=5, photons=1 << 10, width=6,
                                        height=6), ["T", "TT"]),
}


class GpuCase:
    def __init__(self, scene):
        self.s = scene
        self.g = CA.GpuScene(scene)
        self.nodes = self.g.bvh.download_nodes()
        self.root = self.g.bvh.download_root()
        self.d_nb = self.g.bvh.normal_bounds(self.g.d_attribs.ptr)
        self.nb = oracle.normal_bounds(self.nodes, scene.attribs)
        self.eta = CA.eta_range(scene)
        self.pc = CA.PathCuts()

    def gpu(self, event, capacity=None, desc=None):
        desc = desc or self.g.desc()
        if capacity is None:
            _, n = self.pc.enumerate(desc, self.d_nb, event, *self.eta, capacity=0)
            capacity = max(n, 1)
        d, n = self.pc.enumerate(desc, self.d_nb, event, *self.eta, capacity=capacity)
        assert self.pc.status() == 0
        return CA.download_paths(d, min(n, capacity)), n

    def oracle(self, event, points=None):
        pts = self.s.points if points is None else points
        return oracle.path_cuts(self.s.tris, self.s.attribs, self.nodes, self.root, self.nb, pts, self.s.p_light,
                                event, *self.eta)


_cases = {}


def _case(name):
    if name not in _cases:
        print(f"path cuts scene {name}", flush=True)
        _cases[name] = GpuCase(SMALL[name][0]())
    return _cases[name]


def _expected_records(counts, tuples, event):
    K = event >> 24
    want = np.zeros(tuples.shape[0], dtype=CA.PATH)
    want["pixel"] = np.repeat(np.arange(counts.shape[0], dtype=np.uint32), counts)
    want["event"] = event
    want["photon"] = 0xFFFFFFFF
    want["tris"][:] = -1
    want["tris"][:, :K] = tuples
    want["params"][:, :2 * K] = np.float32(1.0) / np.float32(3.0)
    return want


@pytest.mark.parametrize("name", sorted(SMALL))
def test_records_match_the_oracle(name):
    c = _case(name)
    assert np.array_equal(c.nb.view(np.uint8), c.d_nb.download(np.uint8, c.nb.nbytes))
    found = 0
    for ev in SMALL[name][1]:
        event = CA.event_u32(ev)
        got, n = c.gpu(event)
        counts, tuples = c.oracle(event)
        assert n == tuples.shape[0], f"{name} {ev}: {n} tuples vs the oracle's {tuples.shape[0]}"
        want = _expected_records(counts, tuples, event)
        assert got.tobytes() == want.tobytes(), f"{name} {ev}: records differ"
        found += n
    assert found > 0


def test_capacity_and_count_only():
    c = _case("slab")
    event = CA.event_u32("TT")
    full, n = c.gpu(event)
    assert n > 8
    _, n0 = c.pc.enumerate(c.g.desc(), c.d_nb, event, *c.eta, capacity=0)
    assert n0 == n
    part, n2 = c.gpu(event, capacity=n // 3)
    assert n2 == n and part.tobytes() == full[:n // 3].tobytes()


def test_point_subset_and_missed_points():
    c = _case("pool")
    pts = c.s.points[::2].copy()
    pts["R"][::3] = 0.0  # no diffuse hit: nothing to enumerate
    d_pts = DeviceBuffer.from_array(pts)
    desc = c.g.desc()
    desc.points, desc.n_points = d_pts.ptr, pts.shape[0]
    event = CA.event_u32("TT")
    got, n = c.gpu(event, desc=desc)
    counts, tuples = c.oracle(event, pts)
    assert got.tobytes() == _expected_records(counts, tuples, event).tobytes()
    assert np.all(counts[::3] == 0) and n > 0


def test_many_points_walk_without_expansion():
    # more points than the expansion targets: each point's walk starts at the
    # root; tiled copies of the receiver get the records of their original
    c = _case("slab")
    base = c.s.points
    reps = (1 << 20) // base.shape[0] + 3
    pts = np.tile(base, reps)
    d_pts = DeviceBuffer.from_array(pts)
    desc = c.g.desc()
    desc.points, desc.n_points = d_pts.ptr, pts.shape[0]
    event = CA.event_u32("T")  # two tuples a point: a small output
    got, n = c.gpu(event, desc=desc)
    counts, tuples = c.oracle(event)
    assert n == tuples.shape[0] * reps
    want = _expected_records(counts, tuples, event)
    got = got.reshape(reps, -1)
    for r in (0, 1, reps // 2, reps - 1):
        block = got[r].copy()
        assert np.all(block["pixel"] // base.shape[0] == r)
        block["pixel"] -= np.uint32(r * base.shape[0])
        assert block.tobytes() == want.tobytes(), f"copy {r} differs"


@pytest.fixture(scope="module")
def slab_run():
    s = scenes.config2(cells=16, width=16, photons=1 << 14, n_lambdas=2)
    print(f"path cuts vs photons: {s.n_tris} tris", flush=True)
    c = GpuCase(s)
    c.g.trace()
    c.g.build_table()
    return c


def _keys(pixel, t0, t1):
    return (pixel.astype(np.int64) << 40) | (t0.astype(np.int64) << 20) | t1.astype(np.int64)


def test_contributing_photon_tuples_survive_the_cuts(slab_run):
    # the photon pipeline's TT paths that contribute at some wavelength are
    # among the enumerated tuples of their point
    c = slab_run
    event = CA.event_u32("TT")
    d_paths, n = c.g.lookup(event, capacity=1 << 20)
    paths = CA.download_paths(d_paths, n)
    L = c.s.lambdas.shape[0]
    bs = CA.BatchedSolver(c.g.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res = DeviceBuffer(max(n * L, 1) * CA.PATH_RESULT.itemsize)
    d_res.zero()
    c.g.clear_image()
    bs.solve(CA.upload_scene_descs([c.g.desc()]), 1, L, CA.REFINE_NONE, d_res)
    res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
    res = np.empty_like(res_sorted)
    res[order] = res_sorted
    contributing = (res["status"] == CA.PATH_CONTRIBUTES).any(axis=1)
    assert contributing.sum() > 50
    got, m = c.gpu(event)
    cut = np.unique(_keys(got["pixel"], got["tris"][:, 0], got["tris"][:, 1]))
    pc = paths[contributing]
    photon = np.unique(_keys(pc["pixel"], pc["tris"][:, 0], pc["tris"][:, 1]))
    missing = photon[~np.isin(photon, cut)]
    assert missing.size == 0, f"{missing.size} of {photon.size} contributing photon tuples were cut"
    print(f"path cuts: {m} tuples for {c.s.points.shape[0]} points; {photon.size} photon tuples contribute")


def test_slab_records_match_the_oracle_on_a_point_sample(slab_run):
    c = slab_run
    pts = c.s.points[np.random.default_rng(7).choice(c.s.points.shape[0], 2, replace=False)]
    d_pts = DeviceBuffer.from_array(pts)
    desc = c.g.desc()
    desc.points, desc.n_points = d_pts.ptr, pts.shape[0]
    event = CA.event_u32("TT")
    got, n = c.gpu(event, desc=desc)
    counts, tuples = c.oracle(event, pts)
    assert n > 0 and got.tobytes() == _expected_records(counts, tuples, event).tobytes()


def test_batched_solve_of_enumerated_paths_matches_the_oracle():
    c = _case("mirror_slab")
    event = CA.event_u32("RTT")
    got, n = c.gpu(event)
    d_paths = DeviceBuffer.from_array(got)
    L = c.s.lambdas.shape[0]
    bs = CA.BatchedSolver(c.g.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res = DeviceBuffer(max(n * L, 1) * CA.PATH_RESULT.itemsize)
    d_res.zero()
    c.g.clear_image()
    bs.solve(CA.upload_scene_descs([c.g.desc()]), 1, L, CA.REFINE_NONE, d_res)
    res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
    res = np.empty_like(res_sorted)
    res[order] = res_sorted
    res = res.reshape(-1)
    hs = oracle.HostScene(c.s, c.nodes, c.root)
    want = oracle.solve_paths(hs, got, threads=8)
    for f in ("status", "iterations", "params", "footprint", "contribution"):
        a, b = res[f].reshape(res.shape[0], -1), want[f].reshape(want.shape[0], -1)
        same = np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1)
        assert same.all(), f"field {f} differs in {(~same).sum()} items"
    assert (want["status"] == CA.PATH_CONTRIBUTES).sum() > 0
