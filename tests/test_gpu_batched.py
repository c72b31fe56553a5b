"""GPU tests of the batched solver (palz_batched_*): the chain-length
partition, per-item parity with the oracle's FP32 solve (bit-exact without the
refinement step), the FP64 refinement step against the oracle's FP64 Newton,
hipGraph replays, the autotuner and a multi-scene batch."""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd import types as T
from palzilla_amd.device import DeviceBuffer, download_ptr

pytestmark = pytest.mark.gpu

FIELDS = ("status", "iterations", "params", "footprint", "contribution")


def _scenes():
    return {
        "c2_small": scenes.config2(cells=32, width=96, photons=1 << 14, n_lambdas=4),
        "mirror_slab": scenes.stack_scene(mirror_cells=12, layer="slab", layer_cells=20, ior=1.6, seed=6,
                                          photons=1 << 13, width=48),
        "pool": scenes.stack_scene(layer="pool", layer_cells=24, ior=1.33, seed=5, photons=1 << 13, width=48),
    }


SCENES = _scenes()


def _eta(scene, l):
    lam = np.float32(scene.lambdas[l])
    return float(np.float32(scene.cauchy_a) + np.float32(scene.cauchy_b) / (lam * lam))


def _unsort(sorted_items, order, L):
    """Results of the length-sorted path list back in the caller's path order."""
    shape = (len(order), L) + sorted_items.shape[1:]
    out = np.empty_like(sorted_items)
    out.reshape(shape)[order] = sorted_items.reshape(shape)
    return out


def _vertices(scene, path, params):
    K = int(path["event"]) >> 24
    v = scene.tris[path["tris"][:K]].astype(np.float64).reshape(K, 3, 3)
    p = np.asarray(params, dtype=np.float64)[:2 * K].reshape(K, 2)
    return v[:, 0] + (v[:, 1] - v[:, 0]) * p[:, :1] + (v[:, 2] - v[:, 0]) * p[:, 1:]


class Run:
    def __init__(self, name):
        s = SCENES[name]
        print(f"batched scene {name}: {s.n_tris} tris", flush=True)
        self.name, self.s = name, s
        self.g = CA.GpuScene(s)
        self.g.trace()
        self.g.build_table()
        self.nodes = self.g.bvh.download_nodes()
        self.root = self.g.bvh.download_root()
        self.d_paths, self.n = self.g.lookup(0, capacity=1 << 20)
        self.paths = CA.download_paths(self.d_paths, self.n)
        self.L = s.lambdas.shape[0]
        self.d_scenes = CA.upload_scene_descs([self.g.desc()])
        self.d_scenes_narrow = CA.upload_scene_descs([self.g.desc(wide=False)])
        self.bs = CA.BatchedSolver(self.g.sorter)
        self.bs.prepare(self.d_paths, self.n)
        self.order = self.bs.order()
        self.d_res = DeviceBuffer(max(self.n * self.L, 1) * CA.PATH_RESULT.itemsize)
        self.d_ref = DeviceBuffer(max(self.n * self.L, 1) * 64)
        self._oracle = None

    def solve(self, refine, graph=False, wide=True):
        """One solve into the run's fixed result buffers (cleared first, eagerly),
        so that repeated graph solves replay one captured graph.  wide=False
        solves with a descriptor without the BVH4 view."""
        n, L = self.n, self.L
        self.d_res.zero()
        self.d_ref.zero()
        self.g.clear_image()
        self.bs.solve(self.d_scenes if wide else self.d_scenes_narrow, 1, L, refine, self.d_res, self.d_ref,
                      graph=graph)
        res = _unsort(CA.download_results(self.d_res, n * L), self.order, L)
        ref = _unsort(self.d_ref.download(np.float64, n * L * 8).reshape(n * L, 8), self.order, L)
        return res, ref, self.g.image().reshape(-1)

    def oracle(self):
        """FP32 oracle results of the K <= 3 paths and their image."""
        if self._oracle is None:
            keep = (self.paths["event"] >> 24) <= 3
            hs = oracle.HostScene(self.s, self.nodes, self.root)
            want = oracle.solve_paths(hs, self.paths[keep], threads=8)
            self._oracle = keep, want, hs.image.copy()
        return self._oracle


RUNS = {}


@pytest.fixture(scope="module", params=sorted(SCENES))
def run(request):
    if request.param not in RUNS:
        RUNS[request.param] = Run(request.param)
    return RUNS[request.param]


def test_partition_by_chain_length(run):
    r = run.bs.ranges()
    K = run.paths["event"] >> 24
    assert r[5] == run.n
    for k in range(5):
        assert r[k] == int((K <= k).sum())
    sp = run.bs.sorted_paths()
    assert sp.tobytes() == run.paths[run.order].tobytes()
    ks = sp["event"] >> 24
    assert np.all(np.diff(ks.astype(np.int64)) >= 0), "chain lengths must come in order"
    # within a chain length: by (scene, first triangle), ties in input order
    key = sp["scene"].astype(np.int64) << 32 | sp["tris"][:, 0].astype(np.int64)
    for k in range(1, 5):
        m = ks == k
        kk, idx = key[m], run.order[m].astype(np.int64)
        step = np.diff(kk)
        assert np.all(step >= 0), "paths of one chain length must be ordered by scene and first triangle"
        assert np.all(np.diff(idx)[step == 0] > 0), "equal keys must keep their input order"


def test_batched_matches_oracle_bit_exact(run):
    res, _, img = run.solve(refine=CA.REFINE_NONE)
    keep, want, want_img = run.oracle()
    L = run.L
    keep_items = np.repeat(keep, L)
    got = res[keep_items]
    for f in FIELDS:
        a, b = got[f].reshape(got.shape[0], -1), want[f].reshape(want.shape[0], -1)
        bad = np.nonzero(~np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1))[0]
        assert bad.size == 0, f"{run.name}: field {f} differs in {bad.size}/{len(a)} items; first {bad[:3]}"
    skipped = res[~keep_items]
    assert np.all(skipped["status"] == 0) and np.all(skipped["params"] == 0)
    assert (want["status"] == CA.PATH_CONTRIBUTES).sum() > 0
    assert np.allclose(img, want_img, rtol=1e-5, atol=1e-6)


def test_fixed_point_image_is_exact_and_order_independent(run):
    # 32.32 fixed-point sums: integer additions, so the GPU image equals the
    # oracle's bit for bit, solve after solve (the float image only to rtol)
    n_img = run.s.points.shape[0] * run.L
    keep, want, want_img = run.oracle()
    hs = oracle.HostScene(run.s, run.nodes, run.root, fixed=True)
    oracle.solve_paths(hs, run.paths[keep], threads=8, want_results=False)
    d_fixed = DeviceBuffer(n_img * 8)
    d_scenes = CA.upload_scene_descs([run.g.desc(image_fixed=d_fixed.ptr)])
    images = []
    for graph in (False, True):
        d_fixed.zero()
        run.bs.solve(d_scenes, 1, run.L, CA.REFINE_NONE, run.d_res, graph=graph)
        images.append(d_fixed.download(np.uint64, n_img))
    assert images[0].tobytes() == hs.image_fixed.tobytes()
    assert images[1].tobytes() == images[0].tobytes()
    assert images[0].sum() > 0
    img = CA.image_fixed_to_float(d_fixed, n_img)
    assert img.tobytes() == (images[0].astype(np.float64) * 2.0 ** -32).astype(np.float32).tobytes()
    assert np.allclose(img, want_img, rtol=1e-5, atol=1e-6)


def _wide_boxes(w):
    """(lower, upper) of a wide node's six box slots, each (n, 6, 3), from
    its planes; the third plane row repeats the lower bounds."""
    pl = w["plane"]  # (n, axis, row, slot)
    assert pl[:, :, 2, :].tobytes() == pl[:, :, 0, :].tobytes()
    return np.ascontiguousarray(pl[:, :, 0, :].transpose(0, 2, 1)), \
        np.ascontiguousarray(pl[:, :, 1, :].transpose(0, 2, 1))


def test_wide_nodes_mirror_the_tree(run):
    n_int = run.s.n_tris - 1
    w = download_ptr(run.g.wide, T.WIDE_NODE, n_int)
    nodes = run.nodes
    lo, hi = _wide_boxes(w)
    assert np.array_equal(w["child"], nodes["children"])
    assert lo[:, 0:2].tobytes() == np.ascontiguousarray(nodes["aabbs"]["lower"]).tobytes()
    assert hi[:, 0:2].tobytes() == np.ascontiguousarray(nodes["aabbs"]["upper"]).tobytes()
    for c in range(2):
        ch = nodes["children"][:, c]
        internal = (ch & 0x80000000) == 0
        idx = ch[internal]
        s = slice(2 + 2 * c, 4 + 2 * c)
        assert np.array_equal(w["grand"][internal, 2 * c:2 * c + 2], nodes["children"][idx])
        assert lo[internal, s].tobytes() == np.ascontiguousarray(nodes["aabbs"]["lower"][idx]).tobytes()
        assert hi[internal, s].tobytes() == np.ascontiguousarray(nodes["aabbs"]["upper"][idx]).tobytes()
        assert np.all(w["grand"][~internal, 2 * c:2 * c + 2] == 0x7FFFFFFF)
        assert np.all(lo[~internal, s] == 0) and np.all(hi[~internal, s] == 0)


def test_wide_walk_matches_node_walk(run):
    # the occlusion rays walk the BVH4 view or the nodes: same results
    res_w, _, img_w = run.solve(refine=CA.REFINE_NONE)
    res_n, _, img_n = run.solve(refine=CA.REFINE_NONE, wide=False)
    assert res_w.tobytes() == res_n.tobytes()
    # equal contributions; the image sums them with float atomics in any order
    assert np.allclose(img_w, img_n, rtol=1e-5, atol=1e-6)
    assert (res_w["status"] == CA.PATH_CONTRIBUTES).sum() > 0


def test_batched_stats(run):
    res, _, _ = run.solve(refine=CA.REFINE_NONE)
    st = run.bs.stats()
    K = np.repeat(run.paths["event"] >> 24, run.L)
    for k in (1, 2, 3):
        m = K == k
        row = st[f"K{k}"]
        assert row["items"] == int(m.sum())
        assert row["converged"] == int((res["status"][m] >= 1).sum())
        assert row["contributing"] == int((res["status"][m] == 2).sum())
        # one residual evaluation per Newton iteration plus the converged check
        conv = m & (res["status"] >= 1)
        assert row["iterations"] >= int(res["iterations"][conv].sum()) + int(conv.sum())
    assert st["skipped"] == int((K == 4).sum())


def test_refinement_step_vs_fp64_oracle(run):
    # the refinement step equals the oracle's bit for bit and brings the
    # vertices within 1e-5 of the FP64 oracle's converged chain
    res32, _, _ = run.solve(refine=CA.REFINE_NONE)
    res, refined, img = run.solve(refine=CA.REFINE_VERTICES)
    # decisions and irradiance stay on the FP32 chain: only params change
    for f in ("status", "iterations", "contribution", "footprint"):
        assert res[f].tobytes() == res32[f].tobytes(), f
    _, _, want_img = run.oracle()
    assert np.allclose(img, want_img, rtol=1e-5, atol=1e-6)
    conv = np.nonzero(res32["status"] >= 1)[0]
    assert conv.size > 0
    assert np.all(refined[res32["status"] == 0] == 0)
    rng = np.random.default_rng(0)
    sample = conv if conv.size <= 300 else rng.choice(conv, 300, replace=False)
    err32, err64, step_diff = [], [], []
    for it in sample:
        pi, l = divmod(int(it), run.L)
        path = run.paths[pi]
        K = int(path["event"]) >> 24
        p_end = run.s.points[path["pixel"]]["p"]
        eta = _eta(run.s, l)
        p32 = res32["params"][it][:2 * K].astype(np.float64)
        # the oracle's refinement step from the FP32 solution (FP64 residual,
        # FP32 Jacobian, FP64 correction)
        one = oracle.refine_step(run.s, path, p_end, eta, res32["params"][it])
        step_diff.append(np.max(np.abs(one - refined[it][:2 * K])))
        # and converged to FP64 round-off
        ok, _, x64 = oracle.solve_f64(run.s, path, p_end, eta, p32, tol=1e-28, max_iter=64)
        assert ok, "the FP64 oracle must converge from the FP32 solution"
        x = _vertices(run.s, path, refined[it])
        x32 = _vertices(run.s, path, p32)
        scale = np.linalg.norm(x64, axis=1)
        err64.append(np.max(np.linalg.norm(x - x64, axis=1) / scale))
        err32.append(np.max(np.linalg.norm(x32 - x64, axis=1) / scale))
    step_diff, err64, err32 = map(np.asarray, (step_diff, err64, err32))
    print(f"{run.name}: {len(sample)} chains; FP32 vertex err max {err32.max():.3e}, "
          f"after FP64 step max {err64.max():.3e}; GPU-vs-oracle step diff max {step_diff.max():.3e}")
    assert step_diff.max() == 0.0, "the GPU step equals the oracle's bit for bit"
    # One refinement step squares the FP32 error; a chain that the FP32 rule
    # (cost < 1e-10, pk.h:556) stops far from its root on an ill-conditioned
    # Jacobian can stay above 1e-5 (mirror_slab: 1 of 278, 7.6e-4 -> 1.9e-5).
    assert np.mean(err64 <= 1e-5) >= 0.99 and err64.max() <= 1e-4
    assert np.all(err64 <= err32 + 1e-11)
    # the reported float parameters are the refined ones
    assert np.array_equal(res["params"][sample], refined[sample].astype(np.float32))


def test_refined_chain_feeds_footprint(run):
    res1, ref1, _ = run.solve(refine=CA.REFINE_VERTICES)
    res2, ref2, img = run.solve(refine=CA.REFINE_ALL)
    assert ref1.tobytes() == ref2.tobytes()
    assert res1["status"].tobytes() == res2["status"].tobytes()
    assert res1["params"].tobytes() == res2["params"].tobytes()
    _, _, want_img = run.oracle()
    rel_rmse = np.sqrt(np.mean((img - want_img) ** 2)) / max(np.sqrt(np.mean(want_img ** 2)), 1e-30)
    m = res2["status"] == CA.PATH_CONTRIBUTES
    rel = np.abs(res2["footprint"][m] / res1["footprint"][m] - 1.0)
    print(f"{run.name}: refined-chain irradiance rel RMSE vs FP32 oracle {rel_rmse:.3e}; "
          f"footprint rel change median {np.median(rel):.2e} max {rel.max():.2e}")
    assert rel_rmse < 5e-2


def test_graph_replays_match_eager(run):
    eager, eager_ref, eager_img = run.solve(refine=CA.REFINE_VERTICES)
    for rep in range(3):
        res, ref, img = run.solve(refine=CA.REFINE_VERTICES, graph=True)
        assert res.tobytes() == eager.tobytes(), f"replay {rep}"
        assert ref.tobytes() == eager_ref.tobytes(), f"replay {rep}"
        assert np.allclose(img, eager_img, rtol=1e-5, atol=1e-6), f"replay {rep}"
    # a change of arguments captures a new graph
    res, _, _ = run.solve(refine=CA.REFINE_NONE, graph=True)
    res32, _, _ = run.solve(refine=CA.REFINE_NONE)
    assert res.tobytes() == res32.tobytes()


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_branches_match_eager(run, mode):
    # the captured solve as one stream or as three forked branches (one per
    # chain length): every replay equals the eager solve
    eager, eager_ref, eager_img = run.solve(refine=CA.REFINE_VERTICES)
    run.bs.set_branches(mode)
    try:
        for rep in range(3):
            res, ref, img = run.solve(refine=CA.REFINE_VERTICES, graph=True)
            assert res.tobytes() == eager.tobytes(), f"branches={mode}, replay {rep}"
            assert ref.tobytes() == eager_ref.tobytes(), f"branches={mode}, replay {rep}"
            assert np.allclose(img, eager_img, rtol=1e-5, atol=1e-6), f"branches={mode}, replay {rep}"
    finally:
        run.bs.set_branches(-1)


def test_autotune_keeps_results(run):
    before, _, _ = run.solve(refine=CA.REFINE_VERTICES)
    rep = run.bs.autotune(run.d_scenes, 1, run.L, refine=CA.REFINE_VERTICES)
    K = run.paths["event"] >> 24
    for k in (1, 2, 3):
        row = rep[k - 1]
        assert row["block"] in (64, 128, 256) and row["finish_block"] in (64, 128, 256)
        assert 1 <= row["blocks_per_cu"] <= 8 and row["chunk"] in (64, 256, 1024)
        cfg = run.bs.get_config(k)
        assert {f: cfg[f] for f in ("block", "blocks_per_cu", "chunk", "finish_block")} == \
            {f: row[f] for f in ("block", "blocks_per_cu", "chunk", "finish_block")}
        if (K == k).any():
            assert row["newton_us"] > 0
    after, _, _ = run.solve(refine=CA.REFINE_VERTICES)
    assert after.tobytes() == before.tobytes()
    # mode -1: the autotuner timed whole solves per length and fused and kept one
    mode, active = run.bs.get_fused()
    assert mode == -1
    run.bs.set_fused((active + 1) % 4)
    other_mode, _, _ = run.solve(refine=CA.REFINE_VERTICES)
    run.bs.set_fused(-1)
    assert run.bs.get_fused() == (-1, active)
    assert other_mode.tobytes() == before.tobytes()
    run.bs.set_config(1, 64, 1, 1, 64)
    odd, _, _ = run.solve(refine=CA.REFINE_VERTICES)
    assert odd.tobytes() == before.tobytes()


def test_fused_kernels_match_per_length_kernels(run):
    # one persistent Newton and one occlusion kernel over every chain length:
    # the same items, results and refined vertices as one kernel per length,
    # eagerly and in graph replays
    split, split_ref, split_img = run.solve(refine=CA.REFINE_VERTICES)
    try:
        # fused Newton (1), fused occlusion (2), both (3); the fused Newton
        # kernel's register budget: 2 or 3 waves per SIMD
        for mask, waves in ((1, 2), (2, 2), (3, 2), (3, 3)):
            run.bs.set_fused(mask)
            run.bs.set_config(0, 256, 2, waves, 256)
            for graph in (False, True, True):
                res, ref, img = run.solve(refine=CA.REFINE_VERTICES, graph=graph)
                case = f"mask={mask} waves={waves} graph={graph}"
                assert res.tobytes() == split.tobytes(), case
                assert ref.tobytes() == split_ref.tobytes(), case
                assert np.allclose(img, split_img, rtol=1e-5, atol=1e-6), case
        st = run.bs.stats()
        K = run.paths["event"] >> 24
        for k in (1, 2, 3):
            assert st[f"K{k}"]["items"] == int((K == k).sum()) * run.L
    finally:
        run.bs.set_fused(-1)


def test_autotune_tile_shape_cache(run, tmp_path):
    # the first tuning of a workload appends one line to the cache; a second
    # solver on the same workload takes its shapes from there without timing
    cache = tmp_path / "tiles.txt"
    run.bs.set_tune_cache(str(cache))
    try:
        rep = run.bs.autotune(run.d_scenes, 1, run.L, refine=CA.REFINE_VERTICES)
        lines = cache.read_text().splitlines()
        assert len(lines) == 1 and lines[0].startswith("palz-tune ")
        tuned = [run.bs.get_config(k) for k in (1, 2, 3)]
        other = CA.BatchedSolver(run.g.sorter)
        other.prepare(run.d_paths, run.n)
        for k in (1, 2, 3):
            other.set_config(k, 64, 1, 1, 64)
        other.set_tune_cache(str(cache))
        rep2 = other.autotune(run.d_scenes, 1, run.L, refine=CA.REFINE_VERTICES)
        assert rep2 == rep
        assert [other.get_config(k) for k in (1, 2, 3)] == tuned
        assert len(cache.read_text().splitlines()) == 1, "a cache hit appends nothing"
        # fused mode is another workload key: it tunes and appends
        other.set_fused(True)
        other.autotune(run.d_scenes, 1, run.L, refine=CA.REFINE_VERTICES)
        assert len(cache.read_text().splitlines()) == 2
        m = other.get_config(0)
        assert m["block"] in (64, 128, 256) and m["finish_block"] in (64, 128, 256) and m["chunk"] in (2, 3)
        # a line with shapes out of range is not applied
        key = lines[0].split(":")[0]
        cache.write_text(key + ": " + " ".join(["7"] * 23) + "\n")
        other.set_fused(-1)
        other.set_config(1, 64, 1, 1, 64)
        other.autotune(run.d_scenes, 1, run.L, refine=CA.REFINE_VERTICES)
        assert other.get_config(1)["block"] in (64, 128, 256)
        assert len(cache.read_text().splitlines()) == 2
        before, _, _ = run.solve(refine=CA.REFINE_VERTICES)
    finally:
        run.bs.set_tune_cache(None)
    assert before.shape[0] == run.n * run.L


def _c4_small():
    return scenes.config4_scenes(n_scenes=6, seed=11, width=24, photon_exp=(11, 12), cell_exp=(3, 4))


def _c4_real_slice():
    # the first 12 scenes of the 256-scene batched workload, at their real
    # sizes (16^2-256^2 cells, 2^14-2^20 photons, K = 1-3, IOR 1.3-1.8)
    return scenes.config4_scenes()[:12]


@pytest.mark.parametrize("make", [_c4_small, _c4_real_slice], ids=["small", "c4_real_slice"])
def test_multi_scene_batch(make):
    scene_list = make()
    print(f"{len(scene_list)} scenes, {sum(x.n_tris for x in scene_list)} triangles", flush=True)
    batch = CA.SceneBatch(scene_list)
    n = batch.build()
    L = batch.n_lambdas
    assert n == sum(batch.counts) and n > 0
    paths = CA.download_paths(batch.d_paths, n)
    assert np.array_equal(np.unique(paths["scene"]), np.nonzero(batch.counts)[0])
    bs = CA.BatchedSolver(batch.sorter)
    bs.prepare(batch.d_paths, n)
    d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    batch.clear_images()
    bs.solve(batch.d_scenes, len(scene_list), L, False, d_res, None, graph=True)
    res = _unsort(CA.download_results(d_res, n * L), bs.order(), L)
    st = bs.stats()
    assert sum(st[f"K{k}"]["items"] for k in (1, 2, 3)) + st["skipped"] == n * L
    # the fused (one kernel over every chain length) solve of the same batch
    bs.set_fused(True)
    d_res_fused = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    batch.clear_images()
    bs.solve(batch.d_scenes, len(scene_list), L, False, d_res_fused, None, graph=True)
    assert CA.download_results(d_res_fused, n * L).tobytes() == CA.download_results(d_res, n * L).tobytes()
    bs.set_fused(False)
    batch.clear_images()
    bs.solve(batch.d_scenes, len(scene_list), L, False, d_res, None, graph=True)
    compared = 0
    for i, (s, g) in enumerate(zip(scene_list, batch.scenes)):
        m = (paths["scene"] == i) & ((paths["event"] >> 24) <= 3)
        if not m.any():
            continue
        hs = oracle.HostScene(s, g.bvh.download_nodes(), g.bvh.download_root())
        sub = paths[m].copy()
        sub["scene"] = 0
        want = oracle.solve_paths(hs, sub, threads=8)
        got = res[np.repeat(m, L)]
        for f in FIELDS:
            a, b = got[f].reshape(got.shape[0], -1), want[f].reshape(want.shape[0], -1)
            assert np.all((a == b) | (np.isnan(a) & np.isnan(b))), f"scene {i} field {f}"
        assert np.allclose(g.image().reshape(-1), hs.image, rtol=1e-5, atol=1e-6), f"scene {i} image"
        compared += got.shape[0]
    print(f"{n} paths, {compared} items equal to the oracle's", flush=True)
    assert compared > 0


def _eta_range(s):
    lam = s.lambdas.astype(np.float32)
    eta = np.float32(s.cauchy_a) + np.float32(s.cauchy_b) / (lam * lam)
    return float(eta.min()), float(eta.max())


@pytest.mark.parametrize("far", [False, True])
def test_interval_cull_matches_oracle(run, far):
    paths = run.paths.copy()
    if far:  # receiver points far from the tuples' caustics
        n_pts = run.s.points.shape[0]
        paths["pixel"] = (paths["pixel"] + n_pts // 2 + run.s.width // 2) % n_pts
    d_paths = DeviceBuffer.from_array(paths)
    got = CA.interval_cull(run.d_scenes, 1, d_paths, paths.shape[0])
    want = oracle.interval_cull(run.s, paths, *_eta_range(run.s))
    K = paths["event"] >> 24
    want[K > 3] = 1
    assert np.array_equal(got, want), f"{np.sum(got != want)} of {len(got)} flags differ"
    print(f"{run.name} far={far}: culled {np.sum(got == 0)} of {len(got)}")


@pytest.mark.parametrize("far", [False, True])
def test_affine_cull_keeps_a_subset_of_the_interval_keeps(run, far):
    """The affine test runs on what the interval test keeps, so it keeps a
    subset of it; no tuple it cuts holds a path the solver makes contribute."""
    paths = run.paths.copy()
    if far:
        n_pts = run.s.points.shape[0]
        paths["pixel"] = (paths["pixel"] + n_pts // 2 + run.s.width // 2) % n_pts
    d_paths = DeviceBuffer.from_array(paths)
    iv_keep = CA.interval_cull(run.d_scenes, 1, d_paths, paths.shape[0])
    af_keep = CA.affine_cull(run.d_scenes, 1, d_paths, paths.shape[0])
    assert np.all(af_keep <= iv_keep)
    K = paths["event"] >> 24
    assert np.all(af_keep[K > 3] == 1)
    if not far:
        keep, want, _ = run.oracle()
        contributes = (want["status"] == CA.PATH_CONTRIBUTES).reshape(-1, run.L).any(axis=1)
        assert not np.any(contributes & (af_keep[keep] == 0)), "a contributing path was culled"
    print(f"{run.name} far={far}: interval keeps {int(iv_keep.sum())}, affine keeps {int(af_keep.sum())} "
          f"of {len(af_keep)}")


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_culling_keeps_every_contribution(run, mode):
    run.bs.set_culling(mode)
    try:
        res, _, img = run.solve(refine=CA.REFINE_NONE)
        st = run.bs.stats()
    finally:
        run.bs.set_culling(False)
    keep, want, want_img = run.oracle()
    got = res[np.repeat(keep, run.L)]
    culled = got["status"] == CA.PATH_CULLED
    assert st["culled"] == int(culled.sum())
    assert not np.any(want["status"][culled] == CA.PATH_CONTRIBUTES)
    for f in FIELDS:
        a, b = got[f][~culled], want[f][~culled]
        assert a.tobytes() == b.tobytes(), f
    assert np.allclose(img, want_img, rtol=1e-5, atol=1e-6)
