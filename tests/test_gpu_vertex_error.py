"""Vertex error of the batched solver against the FP64 oracle, per
configuration: the FP32 Newton solution (the reference's stopping rule,
cost < 1e-10) and the solution after the one FP64 refinement step, on up to
256 converged chains of each of C1 (full size), the first 12 scenes of C4 (real
sizes) and C5 (full size).  The error of a chain is the largest distance of
its vertices from those of the FP64 oracle's converged chain, relative to the
diagonal of the scene's specular geometry (scenes.vertex_error_scale), so it
does not depend on where the scene sits in world space."""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu


def _configs():
    return {
        "c1": lambda: [scenes.config1()],
        "c4_slice": lambda: scenes.config4_scenes()[:12],
        "c5": lambda: [scenes.config5(scenes.slab_heights(128, 5))],
    }


def _solve(scene_list):
    batch = CA.SceneBatch(scene_list)
    n = batch.build()
    L = batch.n_lambdas
    paths = CA.download_paths(batch.d_paths, n)
    bs = CA.BatchedSolver(batch.sorter)
    bs.prepare(batch.d_paths, n)
    d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    out = {}
    for mode in (CA.REFINE_NONE, CA.REFINE_VERTICES):
        batch.clear_images()
        bs.solve(batch.d_scenes, len(scene_list), L, mode, d_res)
        res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
        res = np.empty_like(res_sorted)
        res[bs.order()] = res_sorted
        out[mode] = res.reshape(-1)
    return paths, out, L


def _vertex_errors(scene_list, paths, r32, r64, L, n_pick=256):
    K_all = np.repeat(paths["event"] >> 24, L)
    conv = np.nonzero((r32["status"] >= 1) & (K_all <= 3))[0]
    rng = np.random.default_rng(5)
    pick = conv if conv.size <= n_pick else rng.choice(conv, n_pick, replace=False)
    e32, e64, ks, failed = [], [], [], 0
    for it in pick:
        pi, l = divmod(int(it), L)
        path = paths[pi]
        s = scene_list[int(path["scene"])]
        K = int(path["event"]) >> 24
        lam = np.float32(s.lambdas[l])
        eta = float(np.float32(s.cauchy_a) + np.float32(s.cauchy_b) / (lam * lam))
        p_end = s.points[path["pixel"]]["p"]
        ok, _, x64 = oracle.solve_f64(s, path, p_end, eta, r32["params"][it][:2 * K].astype(np.float64),
                                      tol=1e-28, max_iter=64)
        if not ok:
            failed += 1
            continue
        v = s.tris[path["tris"][:K]].astype(np.float64).reshape(K, 3, 3)

        def verts(prm):
            p = np.asarray(prm, dtype=np.float64)[:2 * K].reshape(K, 2)
            return v[:, 0] + (v[:, 1] - v[:, 0]) * p[:, :1] + (v[:, 2] - v[:, 0]) * p[:, 1:]

        sc = scenes.vertex_error_scale(s)
        e32.append(np.max(np.linalg.norm(verts(r32["params"][it]) - x64, axis=1) / sc))
        e64.append(np.max(np.linalg.norm(verts(r64["params"][it]) - x64, axis=1) / sc))
        ks.append(K)
    return np.asarray(e32), np.asarray(e64), np.asarray(ks), failed


@pytest.mark.parametrize("name", sorted(_configs()))
def test_vertex_error_per_config(name):
    scene_list = _configs()[name]()
    print(f"{name}: {len(scene_list)} scene(s), {sum(x.n_tris for x in scene_list)} triangles", flush=True)
    paths, out, L = _solve(scene_list)
    r32, r64 = out[CA.REFINE_NONE], out[CA.REFINE_VERTICES]
    # the FP64 step changes params only: decisions and irradiance are the FP32 chain's
    for f in ("status", "iterations", "contribution"):
        assert np.array_equal(r32[f], r64[f]), f
    e32, e64, ks, failed = _vertex_errors(scene_list, paths, r32, r64, L)
    assert e32.size > 50
    for K in sorted(set(ks.tolist())):
        m = ks == K
        print(f"{name} K={K}: {int(m.sum())} chains; FP32 max {e32[m].max():.2e}, p99 "
              f"{np.percentile(e32[m], 99):.2e}, within 1e-5 {np.mean(e32[m] <= 1e-5):.3f}; after the FP64 step "
              f"max {e64[m].max():.2e}, within 1e-5 {np.mean(e64[m] <= 1e-5):.3f}", flush=True)
    print(f"{name}: FP64 oracle did not converge on {failed} of {failed + e32.size}", flush=True)
    assert failed <= 2
    # one FP64 step takes every sampled chain within 1e-5 (measured: at most
    # 5.4e-7 over the three configurations; FP32 alone reaches 1.9e-4)
    assert np.all(e64 <= 1e-5), f"{name}: refined vertex error up to {e64.max():.2e}"
