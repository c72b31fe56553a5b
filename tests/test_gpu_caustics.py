"""GPU parity of the caustic path against the CPU oracle: photon records,
the tuple set and path list, and the plain solver (bit-exact per item), on
reduced versions of the mirror (C1) and slab (C2) configurations and on
mixed-chain stack scenes."""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def _scenes():
    return {
        "c1_small": scenes.config1(width=64),
        "c2_small": scenes.config2(cells=32, width=96, photons=1 << 14, n_lambdas=4),
        "pool": scenes.stack_scene(layer="pool", layer_cells=24, ior=1.33, seed=5, photons=1 << 13, width=48),
        "mirror_slab": scenes.stack_scene(mirror_cells=12, layer="slab", layer_cells=20, ior=1.6, seed=6,
                                          photons=1 << 13, width=48),
    }


SCENES = _scenes()


def _canon(paths):
    order = np.lexsort((paths["photon"], paths["tris"][:, 3], paths["tris"][:, 2], paths["tris"][:, 1],
                        paths["tris"][:, 0], paths["event"], paths["pixel"]))
    return paths[order]


@pytest.fixture(scope="module", params=sorted(SCENES))
def pipeline(request):
    name = request.param
    s = SCENES[name]
    print(f"scene {name}: {s.n_tris} tris, {s.photons_per_triangle} photons/tri", flush=True)
    g = CA.GpuScene(s)
    nodes = g.bvh.download_nodes()
    root = g.bvh.download_root()
    return name, s, g, nodes, root


def test_bvh_matches_oracle(pipeline):
    name, s, g, nodes, root = pipeline
    o_nodes, o_root, *_ = oracle.lbvh(s.tris, bits=32)
    assert root == o_root and nodes.tobytes() == o_nodes.tobytes()


def test_photon_records_bit_exact(pipeline):
    name, s, g, nodes, root = pipeline
    g.trace()
    assert g.trace_status() == 0, "traversal stack overflow"
    got = g.records()
    want, depth = oracle.photon_trace(s, nodes, root)
    assert depth <= 64
    assert (want["event"] != 0).sum() > 100
    assert got.tobytes() == want.tobytes()


def test_tuple_table_and_paths(pipeline):
    name, s, g, nodes, root = pipeline
    g.trace()
    info = g.build_table()
    recs = oracle.photon_trace(s, nodes, root)[0]
    assert info["n_valid"] == int((recs["event"] != 0).sum())
    events = sorted(set(int(e) for e in recs["event"] if e))
    for ev in [0] + events:
        want = oracle.lookup_paths(recs, s.points, event=ev)
        d_paths, n = g.lookup(ev, capacity=max(len(want), 1) + 16)
        assert n == len(want), f"event {CA.event_str(ev) if ev else 'all'}"
        got = CA.download_paths(d_paths, n)
        assert _canon(got).tobytes() == _canon(want).tobytes()
        # grouped by (chain length, tuple): equal tuples are contiguous
        key = [tuple(p["tris"]) + (int(p["event"]),) for p in got]
        seen, prev = set(), None
        for k in key:
            if k != prev:
                assert k not in seen, "paths of one tuple must be contiguous"
                seen.add(k)
                prev = k


def test_lookup_capacity_overflow_reports_count(pipeline):
    name, s, g, nodes, root = pipeline
    g.trace()
    g.build_table()
    _, n_full = g.lookup(0, capacity=1 << 20)
    d_paths, n = g.lookup(0, capacity=max(n_full // 3, 1))
    assert n == n_full


@pytest.mark.parametrize("K", [1, 2, 3, 4])
def test_plain_solver_bit_exact(pipeline, K):
    name, s, g, nodes, root = pipeline
    g.trace()
    g.build_table()
    recs = oracle.photon_trace(s, nodes, root)[0]
    hs = oracle.HostScene(s, nodes, root)
    g.clear_image()
    L = s.lambdas.shape[0]
    total = 0
    events = [ev for ev in sorted(set(int(e) for e in recs["event"] if e)) if ev >> 24 == K]
    if not events:
        pytest.skip(f"{name} has no chains of length {K}")
    for ev in events:
        print(f"{name} {CA.event_str(ev)}", flush=True)
        d_paths, n = g.lookup(ev, capacity=1 << 20)
        if n == 0:
            continue
        paths = CA.download_paths(d_paths, n)
        d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
        g.solve_plain(K, ev, d_paths, n, d_res)
        got = CA.download_results(d_res, n * L)
        want = oracle.solve_paths(hs, paths, threads=8)
        for f in ("status", "iterations", "params", "footprint", "contribution"):
            a, b = got[f].reshape(n * L, -1), want[f].reshape(n * L, -1)
            bad = np.nonzero(~np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1))[0]
            assert bad.size == 0, (f"{name} {CA.event_str(ev)}: field {f} differs in {bad.size}/{n * L} items; "
                                   f"first {bad[:3]}: gpu {got[bad[:3]]} oracle {want[bad[:3]]}")
        total += n * L
    if total == 0:
        pytest.skip(f"{name}: no receiver point sees a chain of length {K}")
    img = g.image().reshape(-1)
    ref = hs.image
    assert np.allclose(img, ref, rtol=1e-5, atol=1e-6)
    rel_rmse = np.sqrt(np.mean((img - ref) ** 2)) / max(np.sqrt(np.mean(ref ** 2)), 1e-30)
    assert rel_rmse < 1e-5


def test_plain_solver_fixed_point_image_exact(pipeline):
    # the plain kernel's 32.32 fixed-point image equals the oracle's bit for bit
    name, s, g, nodes, root = pipeline
    g.trace()
    g.build_table()
    recs = oracle.photon_trace(s, nodes, root)[0]
    hs = oracle.HostScene(s, nodes, root, fixed=True)
    L = s.lambdas.shape[0]
    d_fixed = DeviceBuffer(s.points.shape[0] * L * 8)
    d_fixed.zero()
    for ev in sorted(set(int(e) for e in recs["event"] if e)):
        d_paths, n = g.lookup(ev, capacity=1 << 20)
        if n == 0:
            continue
        g.solve_plain(ev >> 24, ev, d_paths, n, image_fixed=d_fixed.ptr)
        oracle.solve_paths(hs, CA.download_paths(d_paths, n), threads=8, want_results=False)
    got = d_fixed.download(np.uint64, s.points.shape[0] * L)
    assert got.tobytes() == hs.image_fixed.tobytes() and got.sum() > 0


def test_c1_full_size_matches_the_oracle():
    # the reference's CPU-runnable configuration at its real size (64^2-cell
    # mirror, 8,192 triangles, 256^2 receiver, 2^16 photons, 550 nm): the
    # path set, every result field and the fixed-point image equal the oracle's
    s = scenes.config1()
    print(f"C1: {s.n_tris} tris, {s.points.shape[0]} receiver points", flush=True)
    g = CA.GpuScene(s)
    g.trace()
    g.build_table()
    nodes, root = g.bvh.download_nodes(), g.bvh.download_root()
    recs = oracle.photon_trace(s, nodes, root)[0]
    want_paths = oracle.lookup_paths(recs, s.points)
    d_paths, n = g.lookup(0, capacity=max(len(want_paths), 1) + 16)
    assert n == len(want_paths) and n > 1000
    paths = CA.download_paths(d_paths, n)
    assert _canon(paths).tobytes() == _canon(want_paths).tobytes()
    L = s.lambdas.shape[0]
    d_fixed = DeviceBuffer(s.points.shape[0] * L * 8)
    d_fixed.zero()
    bs = CA.BatchedSolver(g.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    bs.solve(CA.upload_scene_descs([g.desc(image_fixed=d_fixed.ptr)]), 1, L, CA.REFINE_NONE, d_res)
    res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
    res = np.empty_like(res_sorted)
    res[order] = res_sorted
    res = res.reshape(-1)
    # chains of length 4 are not part of a batch (reported skipped)
    keep = np.repeat((paths["event"] >> 24) <= 3, L)
    assert np.all(res["status"][~keep] == CA.PATH_NOT_CONVERGED) and bs.stats()["skipped"] == int((~keep).sum())
    hs = oracle.HostScene(s, nodes, root, fixed=True)
    want = oracle.solve_paths(hs, paths[(paths["event"] >> 24) <= 3], threads=8)
    got = res[keep]
    for f in ("status", "iterations", "params", "footprint", "contribution"):
        a, b = got[f].reshape(got.shape[0], -1), want[f].reshape(want.shape[0], -1)
        same = np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1)
        assert same.all(), f"C1: field {f} differs in {(~same).sum()} items"
    assert (want["status"] == CA.PATH_CONTRIBUTES).sum() > 1000
    assert d_fixed.download(np.uint64, s.points.shape[0] * L).tobytes() == hs.image_fixed.tobytes()


def test_c2_full_size_photons_exact_and_sampled_points_match_the_oracle():
    # the spectral slab at its real size (256^2-cell wavy top, 131,076
    # triangles, 2^22 photons, 16 wavelengths): every photon record bit for
    # bit; then, on every 97th point of the 1024^2 receiver, the path list,
    # every result field and the fixed-point image equal the oracle's
    import dataclasses

    s = scenes.config2()
    print(f"C2: {s.n_tris} tris, {s.photons_per_triangle} photons/tri", flush=True)
    g = CA.GpuScene(s)
    g.trace()
    assert g.trace_status() == 0
    nodes, root = g.bvh.download_nodes(), g.bvh.download_root()
    recs, depth = oracle.photon_trace(s, nodes, root)
    assert depth <= 64 and (recs["event"] != 0).sum() > 1_000_000
    assert g.records().tobytes() == recs.tobytes()
    del g
    sub = dataclasses.replace(s, points=np.ascontiguousarray(s.points[::97]))
    print(f"C2 receiver sample: {sub.points.shape[0]} points", flush=True)
    gs = CA.GpuScene(sub)
    gs.trace()
    gs.build_table()
    assert gs.records().tobytes() == recs.tobytes()
    want_paths = oracle.lookup_paths(recs, sub.points)
    d_paths, n = gs.lookup(0, capacity=len(want_paths) + 16)
    assert n == len(want_paths) and n > 100_000
    paths = CA.download_paths(d_paths, n)
    assert _canon(paths).tobytes() == _canon(want_paths).tobytes()
    L = s.lambdas.shape[0]
    d_fixed = DeviceBuffer(sub.points.shape[0] * L * 8)
    d_fixed.zero()
    bs = CA.BatchedSolver(gs.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    bs.solve(CA.upload_scene_descs([gs.desc(image_fixed=d_fixed.ptr)]), 1, L, CA.REFINE_NONE, d_res)
    res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
    res = np.empty_like(res_sorted)
    res[order] = res_sorted
    res = res.reshape(-1)
    keep = np.repeat((paths["event"] >> 24) <= 3, L)
    assert bs.stats()["skipped"] == int((~keep).sum())
    hs = oracle.HostScene(sub, nodes, root, fixed=True)
    want = oracle.solve_paths(hs, paths[(paths["event"] >> 24) <= 3], threads=16)
    got = res[keep]
    for f in ("status", "iterations", "params", "footprint", "contribution"):
        a, b = got[f].reshape(got.shape[0], -1), want[f].reshape(want.shape[0], -1)
        same = np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1)
        assert same.all(), f"C2: field {f} differs in {(~same).sum()} items"
    assert (want["status"] == CA.PATH_CONTRIBUTES).sum() > 100_000
    assert d_fixed.download(np.uint64, sub.points.shape[0] * L).tobytes() == hs.image_fixed.tobytes()
