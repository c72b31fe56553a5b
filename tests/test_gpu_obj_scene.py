"""A scene from the reference's own mesh (bin/assets/table/table.obj, read by
palz_mesh_load_obj) through the caustic path on the GPU: photon tracing, the
tuple table and the batched solve of every chain of length 1-3, bit-exact
against the oracle's solve of the same paths."""
import os

import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

TABLE = os.path.join(os.path.dirname(__file__), "golden", "table.obj")


def test_table_scene_solve_matches_the_oracle():
    s = scenes.obj_scene(TABLE, photons=1 << 16, width=48, n_lambdas=2)
    print(f"table scene: {s.n_tris} tris", flush=True)
    g = CA.GpuScene(s)
    g.trace()
    g.build_table()
    d_paths, n = g.lookup(0, capacity=1 << 20)
    assert n > 0
    paths = CA.download_paths(d_paths, n)
    L = s.lambdas.shape[0]
    bs = CA.BatchedSolver(g.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res = DeviceBuffer(max(n * L, 1) * CA.PATH_RESULT.itemsize)
    d_res.zero()
    g.clear_image()
    bs.solve(CA.upload_scene_descs([g.desc()]), 1, L, CA.REFINE_NONE, d_res)
    res_sorted = CA.download_results(d_res, n * L).reshape(n, L)
    res = np.empty_like(res_sorted)
    res[order] = res_sorted
    res = res.reshape(-1)
    keep = (paths["event"] >> 24) <= 3
    hs = oracle.HostScene(s, g.bvh.download_nodes(), g.bvh.download_root())
    want = oracle.solve_paths(hs, paths[keep], threads=8)
    got = res[np.repeat(keep, L)]
    for f in ("status", "iterations", "params", "footprint", "contribution"):
        a, b = got[f].reshape(got.shape[0], -1), want[f].reshape(want.shape[0], -1)
        same = np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1)
        assert same.all(), f"field {f} differs in {(~same).sum()} items"
    assert (want["status"] == CA.PATH_CONTRIBUTES).sum() > 0
    assert np.allclose(g.image().reshape(-1), hs.image, rtol=1e-5, atol=1e-6)
