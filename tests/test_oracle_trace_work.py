"""The oracle's photonTrace work counters (the bench's photon-trace roofline
numerator): tuples counted equal the records that carry a tuple, photons
counted equal the light-facing specular triangles times the photons per
triangle, and sampling every n-th triangle counts exactly those triangles'
photons."""
import numpy as np

import oracle
from palzilla_amd import scenes
from palzilla_amd import types as T


def _traced_triangles(s):
    t = s.tris.reshape(-1, 3, 3).astype(np.float32)
    ng = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]).astype(np.float32)
    ns = s.attribs["shading_normals"][:, 0].astype(np.float32)
    ng = ng * np.sum(ng * ns, axis=1, dtype=np.float32)[:, None]
    light = np.asarray(s.p_light, dtype=np.float32)
    facing = ~(np.sum(ng * (light - t[:, 0]), axis=1, dtype=np.float32) < 0)
    return facing & (s.attribs["material"] != T.MATERIAL_DIFFUSE)


def test_trace_work_counts():
    s = scenes.config1(width=48, photons=1 << 13)
    nodes, root, *_ = oracle.lbvh(s.tris, bits=32)
    w = oracle.photon_trace_work(s, nodes, root)
    recs = oracle.photon_trace(s, nodes, root)[0]
    traced = _traced_triangles(s)
    assert w["tuples"] == int((recs["event"] != 0).sum()) > 0
    assert w["photons"] == int(traced.sum()) * s.photons_per_triangle
    # every traced photon's first ray enters the root; every hit reads one attribute record
    assert w["nodes"] >= w["photons"] and w["attributes"] >= w["tuples"]
    w4 = oracle.photon_trace_work(s, nodes, root, stride=4)
    assert w4["photons"] == int(traced[::4].sum()) * s.photons_per_triangle
    assert w4["nodes"] < w["nodes"]
