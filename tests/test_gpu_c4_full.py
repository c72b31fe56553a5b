"""The multi-scene batch at its full size: all 256 scenes of the batched-solve
workload (16^2 .. 256^2 cells, chain lengths 1-3, IOR U[1.3, 1.8], 2^14 ..
2^20 photons per scene) solved in one hipGraph launch.

  * every scene's 32.32 fixed-point image equals the image of that scene's
    per-descriptor plain launches (the reference's launch structure,
    pk.h:1480-1586), byte for byte;
  * a stratified sample (the 6 largest scenes, 233^2 .. 249^2 cells and up
    to 155k triangles, and 10 scenes drawn over the rest) equals the
    oracle's FP32 solve in every field of every sampled item.
"""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

FIELDS = ("status", "iterations", "params", "footprint", "contribution")


@pytest.mark.timeout(600)
def test_all_256_scenes_in_one_graph_launch():
    scene_list = scenes.config4_scenes()
    print(f"C4: {len(scene_list)} scenes, {sum(s.n_tris for s in scene_list)} triangles", flush=True)
    batch = CA.SceneBatch(scene_list)
    n = batch.build()
    L = batch.n_lambdas
    paths_all = CA.download_paths(batch.d_paths, n)
    print(f"C4: {n} paths", flush=True)

    bs = CA.BatchedSolver(batch.sorter)
    bs.prepare(batch.d_paths, n)
    order = bs.order()
    fixed_b = [DeviceBuffer(max(g.n_points * L, 1) * 8) for g in batch.scenes]
    for b in fixed_b:
        b.zero()
    d_fx = CA.upload_scene_descs([g.desc(image_fixed=b.ptr) for g, b in zip(batch.scenes, fixed_b)])
    d_res = DeviceBuffer(max(n * L, 1) * CA.PATH_RESULT.itemsize)
    d_res.zero()
    bs.solve(d_fx, len(scene_list), L, CA.REFINE_NONE, d_res, graph=True)
    res_sorted = CA.download_results(d_res, n * L)
    res = np.empty_like(res_sorted)
    res.reshape(n, L)[order] = res_sorted.reshape(n, L)
    stats = bs.stats()
    assert stats["bad_scene"] == 0
    print(f"C4 graph solve: {stats}", flush=True)

    # the plain per-scene, per-descriptor launches into their own fixed-point images
    fixed_p = [DeviceBuffer(max(g.n_points * L, 1) * 8) for g in batch.scenes]
    launches = 0
    for i, (gs, b) in enumerate(zip(batch.scenes, fixed_p)):
        b.zero()
        ev_all = gs.records()["event"]
        for ev in sorted(int(v) for v in np.unique(ev_all[ev_all != 0])):
            if ev >> 24 > 3:
                continue
            _, m = gs.lookup(ev, capacity=1, scene_index=i)
            if m:
                d, m = gs.lookup(ev, capacity=m, scene_index=i)
                gs.solve_plain(ev >> 24, ev, d, m, image_fixed=b.ptr)
                launches += 1
    differ = [i for i, (a, b) in enumerate(zip(fixed_b, fixed_p))
              if a.download(np.uint8, a.nbytes).tobytes() != b.download(np.uint8, b.nbytes).tobytes()]
    print(f"C4: {launches} plain launches; scenes whose images differ: {differ}", flush=True)
    assert not differ
    assert sum(int(b.download(np.uint64, b.nbytes // 8).sum() > 0) for b in fixed_b) > 200

    # stratified oracle sample: the 6 largest scenes and 10 drawn over the rest
    sizes = np.array([s.n_tris for s in scene_list])
    largest = list(np.argsort(sizes)[-6:])
    rng = np.random.default_rng(4)
    rest = [i for i in range(len(scene_list)) if i not in largest]
    pick = largest + list(rng.choice(rest, 10, replace=False))
    assert max(max(scene_list[i].info["mirror_cells"], scene_list[i].info["layer_cells"]) for i in largest) >= 200
    checked = 0
    for i in pick:
        s = scene_list[int(i)]
        mine = np.nonzero(paths_all["scene"] == i)[0]
        mine = mine[(paths_all["event"][mine] >> 24) <= 3]
        if mine.size == 0:
            continue
        take = mine[np.linspace(0, mine.size - 1, min(mine.size, 1500)).astype(np.int64)]
        take = np.unique(take)
        g = batch.scenes[int(i)]
        hs = oracle.HostScene(s, g.bvh.download_nodes(), g.bvh.download_root())
        p = paths_all[take].copy()
        p["scene"] = 0  # the oracle solves one scene
        want = oracle.solve_paths(hs, p, threads=8)
        got = res.reshape(n, L)[take].reshape(-1)
        for f in FIELDS:
            a, w = got[f].reshape(len(got), -1), want[f].reshape(len(want), -1)
            same = np.all((a == w) | (np.isnan(a) & np.isnan(w)), axis=1)
            assert same.all(), f"scene {i} ({s.n_tris} tris): field {f} differs in {(~same).sum()} items"
        checked += len(got)
        print(f"scene {i}: {s.n_tris} tris, {len(got)} items equal the oracle", flush=True)
    assert checked > 10000
