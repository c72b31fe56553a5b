"""GPU ray queries (palz_intersect_rays): the reference's three traversals,
intersect_stackfull, intersect_stackfree and its any-hit form
(minimum_lbvh.h:927-1076), equal to the oracle's in every field of every
hit, on a wavy slab, the reference's table mesh and a dense triangle soup."""
import os

import numpy as np
import pytest

import oracle
from palzilla_amd import caustics as CA
from palzilla_amd import scenes
from palzilla_amd import types as T

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _soup_scene():
    s = scenes.config2(cells=8, width=8, photons=1 << 8, n_lambdas=1)
    tris = scenes.triangle_soup(20_000, seed=8, mu=-2.5, sigma=0.4)
    att = np.zeros(tris.shape[0], dtype=T.TRIANGLE_ATTRIB)
    att["shading_normals"][:] = (0.0, 1.0, 0.0)
    s.tris, s.attribs = tris, att
    return s


SCENES = {
    "slab": lambda: scenes.config2(cells=64, width=16, photons=1 << 10, n_lambdas=1),
    "table": lambda: scenes.obj_scene(os.path.join(GOLD, "table.obj"), photons=1 << 10, width=16, n_lambdas=1),
    "soup": _soup_scene,
}


def _rays(tris, n, rng):
    v = tris.reshape(-1, 3)
    lo, hi = v.min(axis=0), v.max(axis=0)
    c, ext = 0.5 * (lo + hi), float(np.max(hi - lo))
    ro = c + rng.normal(size=(n, 3)) * ext
    rd = c + (rng.random((n, 3)) - 0.5) * ext * 1.2 - ro
    return ro.astype(np.float32), rd.astype(np.float32)


@pytest.mark.parametrize("name", sorted(SCENES))
@pytest.mark.parametrize("query,t_max", [(0, 3.402823466e38), (1, 3.402823466e38), (2, 1.0), (2, 3.402823466e38)])
def test_ray_queries_equal_the_oracle(name, query, t_max):
    s = SCENES[name]()
    print(f"rays: {name} ({s.n_tris} tris), query {query}, t_max {t_max:g}", flush=True)
    g = CA.GpuScene(s)
    nodes, root = g.bvh.download_nodes(), g.bvh.download_root()
    rng = np.random.default_rng(query * 10 + len(name))
    ro, rd = _rays(s.tris, 20_000, rng)
    got = g.intersect_rays(ro, rd, query, t_max)
    want = oracle.intersect_rays(s.tris, nodes, root, ro, rd, query, t_max)
    assert got.tobytes() == want.tobytes()
    hit = want["tri"] != 0xFFFFFFFF
    assert 0.05 < hit.mean() < 0.98
