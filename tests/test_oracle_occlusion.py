"""Occlusion inside the oracle's solver, pinned by geometry (CPU only).

A slab scene's photon-made path list is solved twice by the oracle: in the
scene itself and in the same scene with two diffuse occluder squares appended
(one between the light and the slab, one between the slab and the receiver;
every slab triangle keeps its index).  Newton does not look at the BVH, so the
solutions are identical; contributableThroughput's occlusion rays
(pk.h:88-115, :829-840) must reject exactly the contributing chains whose
polyline light -> x1 -> x2 -> receiver point crosses an occluder, decided
here in float64 from the solved vertices.
"""
import numpy as np

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes

# (y, x0, x1, z0, z1): above the slab (top near y = 0.32, light at y = 1.6)
# and between the slab bottom (y = 0.12) and the receiver (y = 0)
OCCLUDERS = [(0.9, -0.35, 0.05, -0.30, 0.20), (0.06, 0.05, 0.55, -0.45, 0.05)]


def occluder_scene_pair():
    base = scenes.stack_scene(layer="slab", layer_cells=14, ior=1.5, seed=6, photons=1 << 13, width=40)
    return base, scenes.with_occluders(base, OCCLUDERS)


def _crossings(poly, margin):
    """Per chain: (crosses an occluder with the given inner margin, passes
    within the margin of an occluder's rim).  poly: (n, V, 3) float64."""
    blocked = np.zeros(poly.shape[0], bool)
    near = np.zeros(poly.shape[0], bool)
    for a, b in zip(poly[:, :-1].transpose(1, 0, 2), poly[:, 1:].transpose(1, 0, 2)):
        for y, x0, x1, z0, z1 in OCCLUDERS:
            dy = b[:, 1] - a[:, 1]
            with np.errstate(divide="ignore", invalid="ignore"):
                t = (y - a[:, 1]) / dy
            ok = (t > 1e-6) & (t < 1 - 1e-6)
            px = a[:, 0] + t * (b[:, 0] - a[:, 0])
            pz = a[:, 2] + t * (b[:, 2] - a[:, 2])
            inside = (px > x0 + margin) & (px < x1 - margin) & (pz > z0 + margin) & (pz < z1 - margin)
            outside = (px < x0 - margin) | (px > x1 + margin) | (pz < z0 - margin) | (pz > z1 + margin)
            blocked |= ok & inside
            near |= ok & ~inside & ~outside
    return blocked, near


def test_solver_rejects_exactly_the_chains_an_occluder_blocks():
    base, occ = occluder_scene_pair()
    n0, r0, *_ = oracle.lbvh(base.tris, bits=32)
    recs, _ = oracle.photon_trace(base, n0, r0)
    paths = oracle.lookup_paths(recs, base.points)
    paths = paths[(paths["event"] >> 24) == 2]
    assert len(paths) > 1500
    want0 = oracle.solve_paths(oracle.HostScene(base, n0, r0), paths, threads=8)
    n1, r1, *_ = oracle.lbvh(occ.tris, bits=32)
    want1 = oracle.solve_paths(oracle.HostScene(occ, n1, r1), paths, threads=8)
    assert np.array_equal(want0["params"].view(np.uint32), want1["params"].view(np.uint32))
    assert np.array_equal(want0["iterations"], want1["iterations"])

    live = want0["status"] == CA.PATH_CONTRIBUTES
    L = base.lambdas.shape[0]
    pth = np.repeat(paths, L)[live]
    prm = want0["params"][live].astype(np.float64)
    verts = []
    for k in range(2):
        v = base.tris[pth["tris"][:, k]].astype(np.float64).reshape(-1, 3, 3)
        u, w = prm[:, 2 * k:2 * k + 1], prm[:, 2 * k + 1:2 * k + 2]
        verts.append(v[:, 0] + (v[:, 1] - v[:, 0]) * u + (v[:, 2] - v[:, 0]) * w)
    light = np.broadcast_to(np.asarray(base.p_light, np.float64), verts[0].shape)
    end = base.points["p"][pth["pixel"]].astype(np.float64)
    blocked, near = _crossings(np.stack([light, verts[0], verts[1], end], axis=1), 1e-4)

    rejected = want1["status"][live] == CA.PATH_CONVERGED
    kept = want1["status"][live] == CA.PATH_CONTRIBUTES
    assert np.all(rejected | kept)
    assert blocked.sum() > 50 and (~blocked & ~near).sum() > 1000
    assert np.all(rejected[blocked]), "a chain through an occluder still contributes"
    assert not np.any(rejected[~blocked & ~near]), "a free chain was rejected"
    # non-contributing items are unchanged by the occluders
    assert np.array_equal(want0["status"][~live], want1["status"][~live])
