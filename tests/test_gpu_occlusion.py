"""GPU occlusion on a scene with deliberate occluders.

The slab scene's photon-made path list (traced and looked up on the GPU) is
solved in the same scene with two diffuse occluder squares appended (see
test_oracle_occlusion.py, whose CPU test pins the oracle's rejections by
geometry).  The batched solver (BVH4 packet walk in k_occlude) and the plain
per-tuple kernel must both equal the oracle in every field, and a non-zero
number of chains that contribute without the occluders must be rejected by
them.
"""
import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd.device import DeviceBuffer
from test_oracle_occlusion import occluder_scene_pair

pytestmark = pytest.mark.gpu

FIELDS = ("status", "iterations", "params", "footprint", "contribution")


def _same(got, want, what):
    for f in FIELDS:
        a, b = got[f].reshape(len(got), -1), want[f].reshape(len(want), -1)
        bad = np.nonzero(~np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1))[0]
        assert bad.size == 0, f"{what}: field {f} differs in {bad.size}/{len(got)} items, first {bad[:3]}"


def test_occluders_reject_chains_identically_on_gpu_and_oracle():
    base, occ = occluder_scene_pair()
    print(f"occlusion: {base.n_tris} + {occ.n_tris - base.n_tris} tris", flush=True)
    g0 = CA.GpuScene(base)
    g0.trace()
    g0.build_table()
    TT = CA.event_u32("TT")
    d_paths, n = g0.lookup(TT, capacity=1 << 18)
    paths = CA.download_paths(d_paths, n)
    L = base.lambdas.shape[0]
    assert n > 1500

    g1 = CA.GpuScene(occ, sorter=g0.sorter)
    nodes1, root1 = g1.bvh.download_nodes(), g1.bvh.download_root()
    hs1 = oracle.HostScene(occ, nodes1, root1)
    want1 = oracle.solve_paths(hs1, paths, threads=8)
    want0 = oracle.solve_paths(oracle.HostScene(base, g0.bvh.download_nodes(), g0.bvh.download_root()), paths,
                               threads=8)
    rejected = (want0["status"] == CA.PATH_CONTRIBUTES) & (want1["status"] == CA.PATH_CONVERGED)
    assert rejected.sum() > 50, "the occluders must reject contributing chains"

    # plain per-tuple kernel
    d_res = DeviceBuffer(n * L * CA.PATH_RESULT.itemsize)
    g1.solve_plain(2, TT, d_paths, n, d_res)
    _same(CA.download_results(d_res, n * L), want1, "plain")

    # batched solver (FP32 only: no refinement step), results back in path order
    bs = CA.BatchedSolver(g0.sorter)
    bs.prepare(d_paths, n)
    order = bs.order()
    d_res.zero()
    g1.clear_image()
    bs.solve(CA.upload_scene_descs([g1.desc()]), 1, L, CA.REFINE_NONE, d_res)
    sorted_res = CA.download_results(d_res, n * L)
    got = np.empty_like(sorted_res)
    got.reshape(n, L)[order] = sorted_res.reshape(n, L)
    _same(got, want1, "batched")
    print(f"occlusion: {int(rejected.sum())} of {int((want0['status'] == CA.PATH_CONTRIBUTES).sum())} "
          f"contributing items rejected by the occluders", flush=True)
