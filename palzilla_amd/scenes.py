"""Synthetic inputs of the benchmark configurations (seeded, float32).

Every generator is deterministic in its seed (numpy PCG64).  Geometry is
computed in float64 and rounded once to float32.

Scene layout (y is up).  A caustic scene is a stack, top to bottom, of an
optional ceiling mirror, a point light, an optional refracting layer (a wavy
"pool" surface whose medium reaches the floor, or a slab with a wavy top and
a flat bottom) and a diffuse receiver plane at y = 0.  The receiver points are
the pixel centres of a W x H grid on that plane.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import types as T


@dataclass
class CausticScene:
    tris: np.ndarray            # (n, 9) float32
    attribs: np.ndarray         # (n,) TRIANGLE_ATTRIB
    p_light: tuple
    points: np.ndarray          # (W*H,) FIRST_DIFFUSE records
    width: int
    height: int
    extent: tuple               # receiver (x0, x1, z0, z1)
    cauchy_a: float
    cauchy_b: float
    lambdas: np.ndarray         # float32 nm
    photons_per_triangle: int
    eta_min: float
    eta_max: float
    name: str = ""
    info: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return self.tris.shape[0]


FIRST_DIFFUSE = np.dtype([("p", "<f4", 3), ("ng", "<f4", 3), ("R", "<f4", 3), ("lambda", "<f4")])
assert FIRST_DIFFUSE.itemsize == 40


def _sines(rng, n_sines: int, amplitude: float, fmin=1.0, fmax=6.0):
    f = rng.uniform(fmin, fmax, n_sines)
    ang = rng.uniform(0.0, 2.0 * np.pi, n_sines)
    d = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    phase = rng.uniform(0.0, 2.0 * np.pi, n_sines)
    return [(amplitude, fi, di, pi) for fi, di, pi in zip(f, d, phase)]


def _height(sines, x, z):
    h = np.zeros_like(x)
    hx = np.zeros_like(x)
    hz = np.zeros_like(x)
    for a, f, d, ph in sines:
        arg = 2.0 * np.pi * f * (d[0] * x + d[1] * z) + ph
        h += a * np.sin(arg)
        c = a * np.cos(arg) * 2.0 * np.pi * f
        hx += c * d[0]
        hz += c * d[1]
    return h, hx, hz


def heightfield(cells: int, y0: float, sines, extent=(-1.0, 1.0, -1.0, 1.0), facing: float = 1.0,
                heights: np.ndarray | None = None):
    """Triangulated heightfield y = y0 + h(x, z) on a cells x cells grid.

    Returns (tris (2*cells^2, 9) float64, shading normals (2*cells^2, 3, 3)
    float64, vertex heights (cells+1)^2).  Shading normals are the analytic
    surface normals, oriented to +y when facing > 0 and to -y otherwise.
    `heights` (optional) replaces y0 + h at the vertices; normals then come
    from central differences of the vertex heights."""
    x0, x1, z0, z1 = extent
    xs = np.linspace(x0, x1, cells + 1)
    zs = np.linspace(z0, z1, cells + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    if heights is None:
        h, hx, hz = _height(sines, X, Z)
        Y = y0 + h
    else:
        Y = heights.reshape(cells + 1, cells + 1)
        hx = np.gradient(Y, xs, axis=0)
        hz = np.gradient(Y, zs, axis=1)
    N = np.stack([-hx, np.ones_like(hx), -hz], axis=-1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    if facing < 0:
        N = -N
    P = np.stack([X, Y, Z], axis=-1)
    a = P[:-1, :-1].reshape(-1, 3)
    b = P[1:, :-1].reshape(-1, 3)
    c = P[1:, 1:].reshape(-1, 3)
    d = P[:-1, 1:].reshape(-1, 3)
    na = N[:-1, :-1].reshape(-1, 3)
    nb = N[1:, :-1].reshape(-1, 3)
    nc = N[1:, 1:].reshape(-1, 3)
    nd = N[:-1, 1:].reshape(-1, 3)
    t1 = np.stack([a, b, c], axis=1)
    t2 = np.stack([a, c, d], axis=1)
    n1 = np.stack([na, nb, nc], axis=1)
    n2 = np.stack([na, nc, nd], axis=1)
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3, 3)
    nrm = np.stack([n1, n2], axis=1).reshape(-1, 3, 3)
    return tris, nrm, Y.reshape(-1)


def quad(y: float, extent, normal_y: float):
    x0, x1, z0, z1 = extent
    a, b, c, d = (x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1)
    tris = np.array([[a, b, c], [a, c, d]], dtype=np.float64)
    nrm = np.zeros((2, 3, 3))
    nrm[..., 1] = normal_y
    return tris, nrm


def receiver_points(width: int, height: int, extent) -> np.ndarray:
    x0, x1, z0, z1 = extent
    xs = x0 + (np.arange(width) + 0.5) * (x1 - x0) / width
    zs = z0 + (np.arange(height) + 0.5) * (z1 - z0) / height
    Zg, Xg = np.meshgrid(zs, xs, indexing="ij")  # row-major: pixel = x + z * width
    pts = np.zeros(width * height, dtype=FIRST_DIFFUSE)
    pts["p"][:, 0] = Xg.reshape(-1)
    pts["p"][:, 1] = 0.0
    pts["p"][:, 2] = Zg.reshape(-1)
    pts["ng"][:, 1] = 1.0
    pts["R"] = 1.0
    pts["lambda"] = 550.0
    return pts


def _assemble(parts):
    tris = np.concatenate([p[0] for p in parts]).reshape(-1, 9).astype(np.float32)
    att = np.zeros(tris.shape[0], dtype=T.TRIANGLE_ATTRIB)
    att["material"] = np.concatenate([np.full(p[0].shape[0], p[2], dtype=np.int32) for p in parts])
    att["shading_normals"] = np.concatenate([p[1] for p in parts]).astype(np.float32)
    return tris, att


def stack_scene(*, mirror_cells=0, layer="none", layer_cells=0, ior=1.5, cauchy_b=0.0, seed=0,
                photons=1 << 16, width=128, height=128, lambdas=(550.0,), mirror_amp=0.02,
                layer_amp=0.03, n_sines_mirror=4, n_sines_layer=6, receiver_extent=(-1.0, 1.0, -1.0, 1.0),
                name="") -> CausticScene:
    """A caustic scene of the stack family (see the module docstring).

    mirror_cells > 0 adds a ceiling mirror at y = 0.6 facing down and puts the
    light at (0, 0.45, 0); otherwise the light is at (0, 1.6, 0).  layer is
    "none", "pool" (wavy surface at y = 0.3, medium down to the floor) or
    "slab" (wavy top at y = 0.32, flat bottom at y = 0.12)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    spec = 0
    if mirror_cells:
        t, n, _ = heightfield(mirror_cells, 0.6, _sines(rng, n_sines_mirror, mirror_amp), facing=-1.0)
        parts.append((t, n, T.MATERIAL_MIRROR))
        spec += t.shape[0]
        p_light = (0.0, 0.45, 0.0)
    else:
        p_light = (0.0, 1.6, 0.0)
    if layer == "pool":
        t, n, _ = heightfield(layer_cells, 0.3, _sines(rng, n_sines_layer, layer_amp), facing=1.0)
        parts.append((t, n, T.MATERIAL_DIELECTRIC))
        spec += t.shape[0]
    elif layer == "slab":
        t, n, _ = heightfield(layer_cells, 0.32, _sines(rng, n_sines_layer, layer_amp), facing=1.0)
        parts.append((t, n, T.MATERIAL_DIELECTRIC))
        spec += t.shape[0]
        tb, nb = quad(0.12, (-1.0, 1.0, -1.0, 1.0), -1.0)
        parts.append((tb, nb, T.MATERIAL_DIELECTRIC))
    tr, nr = quad(0.0, (-3.0, 3.0, -3.0, 3.0), 1.0)
    parts.append((tr, nr, T.MATERIAL_DIFFUSE))
    tris, att = _assemble(parts)
    lam = np.asarray(lambdas, dtype=np.float32)
    ppt = max(1, int(photons) // max(spec, 1))
    n_lo = ior + cauchy_b / float(lam.min()) ** 2
    n_hi = ior + cauchy_b / float(lam.max()) ** 2
    return CausticScene(tris=tris, attribs=att, p_light=p_light,
                        points=receiver_points(width, height, receiver_extent), width=width, height=height,
                        extent=receiver_extent, cauchy_a=ior, cauchy_b=cauchy_b, lambdas=lam,
                        photons_per_triangle=ppt, eta_min=n_lo, eta_max=n_hi, name=name,
                        info={"mirror_cells": mirror_cells, "layer": layer, "layer_cells": layer_cells,
                              "ior": ior, "photons": photons, "seed": seed})


def vertex_error_scale(scene: CausticScene) -> float:
    """The length relative vertex errors are measured against: the diagonal of
    the bounding box of the scene's specular triangles (where chain vertices
    lie).  It moves with the geometry, so the relative error does not depend
    on where the scene sits in world space."""
    spec = scene.attribs["material"] != T.MATERIAL_DIFFUSE
    v = scene.tris[spec].reshape(-1, 3).astype(np.float64)
    return float(np.linalg.norm(v.max(axis=0) - v.min(axis=0)))


def with_occluders(scene: CausticScene, squares) -> CausticScene:
    """The scene with diffuse occluder squares appended after its triangles
    (so every existing triangle keeps its index and a path list made for the
    scene stays valid).  squares: (y, x0, x1, z0, z1) per horizontal square,
    two triangles each, shading normals +y."""
    parts = []
    for y, x0, x1, z0, z1 in squares:
        a, b, c, d = (x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1)
        parts.append([a + b + c, a + c + d])
    extra = np.asarray(parts, np.float64).reshape(-1, 9).astype(np.float32)
    att = np.zeros(extra.shape[0], dtype=T.TRIANGLE_ATTRIB)
    att["material"] = T.MATERIAL_DIFFUSE
    att["shading_normals"] = np.tile(np.float32([0.0, 1.0, 0.0]), (extra.shape[0], 3, 1))
    info = dict(scene.info, occluders=[tuple(q) for q in squares])
    return CausticScene(tris=np.concatenate([scene.tris, extra]), attribs=np.concatenate([scene.attribs, att]),
                        p_light=scene.p_light, points=scene.points, width=scene.width, height=scene.height,
                        extent=scene.extent, cauchy_a=scene.cauchy_a, cauchy_b=scene.cauchy_b, lambdas=scene.lambdas,
                        photons_per_triangle=scene.photons_per_triangle, eta_min=scene.eta_min,
                        eta_max=scene.eta_max, name=scene.name + "+occluders", info=info)


def config1(seed: int = 1, width: int = 256, photons: int = 1 << 16) -> CausticScene:
    """Single-bounce mirror caustic: 64x64-cell mirror heightfield (8,192
    triangles) over [-1,1]^2 at y = 0.5, 4 sinusoids of amplitude 0.02,
    receiver y = 0 at 256x256, 2^16 photons, 550 nm.  The light sits at
    (0, 0.3, 0) between the downward-facing mirror and the receiver (the
    "(0,3,0)" of the configuration puts light and receiver on opposite sides
    of the mirror, where no one-bounce caustic exists)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, n, _ = heightfield(64, 0.5, _sines(rng, 4, 0.02), facing=-1.0)
    tr, nr = quad(0.0, (-3.0, 3.0, -3.0, 3.0), 1.0)
    tris, att = _assemble([(t, n, T.MATERIAL_MIRROR), (tr, nr, T.MATERIAL_DIFFUSE)])
    lam = np.array([550.0], dtype=np.float32)
    return CausticScene(tris=tris, attribs=att, p_light=(0.0, 0.3, 0.0),
                        points=receiver_points(width, width, (-1.0, 1.0, -1.0, 1.0)), width=width, height=width,
                        extent=(-1.0, 1.0, -1.0, 1.0), cauchy_a=1.5, cauchy_b=0.0, lambdas=lam,
                        photons_per_triangle=max(1, photons // t.shape[0]), eta_min=1.5, eta_max=1.5,
                        name="C1 mirror 64x64, 256^2, 2^16 photons", info={"seed": seed})


def config2(seed: int = 2, cells: int = 256, width: int = 1024, photons: int = 1 << 22,
            n_lambdas: int = 16) -> CausticScene:
    """Spectral two-interface refraction: glass slab with a cells x cells wavy
    top (6 sinusoids, amplitude 0.03) at y = 1.0 and a flat bottom (2
    triangles) at y = 0.5 over [-1,1]^2, Cauchy n = 1.5 + 4200 nm^2 / lambda^2,
    n_lambdas wavelengths on [380, 720] nm, light (0, 3, 0), receiver y = 0
    over [-1.25, 1.25]^2 at width^2."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, n, _ = heightfield(cells, 1.0, _sines(rng, 6, 0.03), facing=1.0)
    tb, nb = quad(0.5, (-1.0, 1.0, -1.0, 1.0), -1.0)
    tr, nr = quad(0.0, (-3.0, 3.0, -3.0, 3.0), 1.0)
    tris, att = _assemble([(t, n, T.MATERIAL_DIELECTRIC), (tb, nb, T.MATERIAL_DIELECTRIC),
                           (tr, nr, T.MATERIAL_DIFFUSE)])
    lam = np.linspace(380.0, 720.0, n_lambdas).astype(np.float32)
    a, b = 1.5, 4200.0
    ext = (-1.25, 1.25, -1.25, 1.25)
    return CausticScene(tris=tris, attribs=att, p_light=(0.0, 3.0, 0.0), points=receiver_points(width, width, ext),
                        width=width, height=width, extent=ext, cauchy_a=a, cauchy_b=b, lambdas=lam,
                        photons_per_triangle=max(1, photons // t.shape[0]),
                        eta_min=a + b / 380.0 ** 2, eta_max=a + b / 720.0 ** 2,
                        name=f"C2 slab {cells}^2 TT, {n_lambdas} lambdas, {width}^2, {photons} photons",
                        info={"seed": seed, "cells": cells})


def slab_heights(cells: int, seed: int, y0: float = 1.0, amplitude: float = 0.03) -> np.ndarray:
    """Vertex heights (cells+1)^2 of a wavy slab top: y0 + 6 seeded sinusoids."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = np.linspace(-1.0, 1.0, cells + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    h, _, _ = _height(_sines(rng, 6, amplitude), X, Z)
    return (y0 + h).reshape(-1)


def config5(heights: np.ndarray, cells: int = 128, width: int = 512, photons: int = 1 << 21,
            ior: float = 1.5) -> CausticScene:
    """Adjoint workload: a glass slab whose top is a cells x cells heightfield
    with the given vertex heights ((cells+1)^2 parameters; shading normals from
    central differences of the heights), flat bottom at y = 0.5, IOR 1.5 at one
    wavelength, light (0, 3, 0), receiver y = 0 over [-1.25, 1.25]^2 at width^2.
    The heightfield's triangles come first (2 per cell, cell-major)."""
    t, n, _ = heightfield(cells, 1.0, [], heights=np.asarray(heights, dtype=np.float64))
    tb, nb = quad(0.5, (-1.0, 1.0, -1.0, 1.0), -1.0)
    tr, nr = quad(0.0, (-3.0, 3.0, -3.0, 3.0), 1.0)
    tris, att = _assemble([(t, n, T.MATERIAL_DIELECTRIC), (tb, nb, T.MATERIAL_DIELECTRIC),
                           (tr, nr, T.MATERIAL_DIFFUSE)])
    ext = (-1.25, 1.25, -1.25, 1.25)
    return CausticScene(tris=tris, attribs=att, p_light=(0.0, 3.0, 0.0), points=receiver_points(width, width, ext),
                        width=width, height=width, extent=ext, cauchy_a=ior, cauchy_b=0.0,
                        lambdas=np.array([550.0], dtype=np.float32),
                        photons_per_triangle=max(1, photons // t.shape[0]), eta_min=ior, eta_max=ior,
                        name=f"C5 slab {cells}^2 TT, {width}^2, {photons} photons",
                        info={"cells": cells, "heightfield_tris": t.shape[0]})


def height_param_tables(cells: int, n_tris: int):
    """The parameterisation of config5's heights for the adjoint: per
    triangle its vertices' height indices (-1 outside the heightfield), per
    height the 4 heights its normal's central differences use (-1: none) and
    their coefficients in (dh/dx, dh/dz)."""
    m = cells + 1
    idx = np.arange(m * m, dtype=np.int32).reshape(m, m)
    a = idx[:-1, :-1].reshape(-1)
    b = idx[1:, :-1].reshape(-1)
    c = idx[1:, 1:].reshape(-1)
    d = idx[:-1, 1:].reshape(-1)
    tv = np.full((n_tris, 3), -1, dtype=np.int32)
    hf = np.stack([np.stack([a, b, c], axis=1), np.stack([a, c, d], axis=1)], axis=1).reshape(-1, 3)
    tv[:hf.shape[0]] = hf
    dx = 2.0 / cells
    nb = np.full((m * m, 4), -1, dtype=np.int32)
    coef = np.zeros((m * m, 8), dtype=np.float32)
    for i in range(m):
        for j in range(m):
            h = i * m + j
            # np.gradient: central inside, one-sided at the border (edge_order 1)
            if 0 < i < cells:
                xs = [(i - 1, -0.5 / dx), (i + 1, 0.5 / dx)]
            elif i == 0:
                xs = [(0, -1.0 / dx), (1, 1.0 / dx)]
            else:
                xs = [(cells - 1, -1.0 / dx), (cells, 1.0 / dx)]
            if 0 < j < cells:
                zs = [(j - 1, -0.5 / dx), (j + 1, 0.5 / dx)]
            elif j == 0:
                zs = [(0, -1.0 / dx), (1, 1.0 / dx)]
            else:
                zs = [(cells - 1, -1.0 / dx), (cells, 1.0 / dx)]
            for e, (ii, cx) in enumerate(xs):
                nb[h, e] = ii * m + j
                coef[h, 2 * e] = cx
            for e, (jj, cz) in enumerate(zs):
                nb[h, 2 + e] = i * m + jj
                coef[h, 2 * (2 + e) + 1] = cz
    return tv, nb, coef


def config4_scenes(n_scenes: int = 256, seed: int = 4, width: int = 64, photon_exp=(14, 20),
                   cell_exp=(4, 8)):
    """Batched workload: n_scenes stack scenes, heightfields of round(2^U[4,8])
    cells per side, chain length K ~ U{1,2,3} (K=1: mirror or pool, K=2: slab or
    mirror+pool, K=3: mirror+slab), IOR ~ U[1.3, 1.8], 2^U{14..20} photons."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    for s in range(n_scenes):
        cells = int(round(2.0 ** rng.uniform(*cell_exp)))
        K = int(rng.integers(1, 4))
        ior = float(rng.uniform(1.3, 1.8))
        photons = 1 << int(rng.integers(photon_exp[0], photon_exp[1] + 1))
        coin = bool(rng.integers(0, 2))
        if K == 1:
            kw = dict(mirror_cells=cells) if coin else dict(layer="pool", layer_cells=cells)
        elif K == 2:
            kw = dict(layer="slab", layer_cells=cells) if coin else dict(mirror_cells=max(4, cells // 2),
                                                                          layer="pool", layer_cells=cells)
        else:
            kw = dict(mirror_cells=max(4, cells // 2), layer="slab", layer_cells=cells)
        scenes.append(stack_scene(ior=ior, seed=seed * 1000 + s, photons=photons, width=width, height=width,
                                  name=f"C4[{s}] K={K}", **kw))
    return scenes


def triangle_soup(n: int, seed: int = 3, mu: float = -6.0, sigma: float = 0.5) -> np.ndarray:
    """LBVH workload: n triangles, centroids U[0,1]^3, size lognormal(mu, sigma),
    uniformly random orientation.  Returns (n, 9) float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 3))
    s = rng.lognormal(mu, sigma, n)
    # uniform random rotations from unit quaternions
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    rot = np.empty((n, 3, 3))
    rot[:, 0, 0] = 1 - 2 * (y * y + z * z)
    rot[:, 0, 1] = 2 * (x * y - z * w)
    rot[:, 0, 2] = 2 * (x * z + y * w)
    rot[:, 1, 0] = 2 * (x * y + z * w)
    rot[:, 1, 1] = 1 - 2 * (x * x + z * z)
    rot[:, 1, 2] = 2 * (y * z - x * w)
    rot[:, 2, 0] = 2 * (x * z - y * w)
    rot[:, 2, 1] = 2 * (y * z + x * w)
    rot[:, 2, 2] = 1 - 2 * (x * x + y * y)
    # equilateral triangle with unit edges, centred at the origin, in the xy-plane
    r = 1.0 / np.sqrt(3.0)
    base = np.array([[r, 0.0, 0.0],
                     [-0.5 * r, 0.5, 0.0],
                     [-0.5 * r, -0.5, 0.0]])
    v = np.einsum("nij,kj->nki", rot, base) * s[:, None, None] + c[:, None, :]
    return v.reshape(n, 9).astype(np.float32)


def load_obj_mesh(path: str, default_material: int = T.MATERIAL_DIFFUSE) -> tuple[np.ndarray, np.ndarray]:
    """Triangles (n, 9) float32 and TriangleAttrib records of a Wavefront OBJ
    file, read by the library's loader (palz_mesh_load_obj: loadFrame's rules,
    pk.h:1246-1362, faces fan-triangulated in file order)."""
    import ctypes as C

    from ._lib import call, lib
    h = C.c_void_p()
    call("palz_mesh_load_obj", os.fsencode(path), int(default_material), C.byref(h))
    try:
        n = C.c_uint32()
        call("palz_mesh_triangle_count", h, C.byref(n))
        tris = np.zeros((n.value, 9), dtype=np.float32)
        att = np.zeros(n.value, dtype=T.TRIANGLE_ATTRIB)
        call("palz_mesh_copy", h, tris.ctypes.data_as(C.c_void_p), att.ctypes.data_as(C.c_void_p))
    finally:
        lib().palz_mesh_destroy(h)
    return tris, att


def obj_scene(path: str, *, lift: float = 1.0, photons: int = 1 << 16, width: int = 64, n_lambdas: int = 4,
              cauchy_a: float = 1.5, cauchy_b: float = 4200.0) -> CausticScene:
    """An OBJ mesh as a glass object over a receiver: every face dielectric
    (unless the file names another material), normals from the file, lifted by
    `lift` in y, light (0.3, 3, 0.2), diffuse floor quad at y = 0 over [-3, 3]^2, receiver [-1.25, 1.25]^2 at width^2, Cauchy dispersion on
    n_lambdas wavelengths over [380, 720] nm."""
    t, a = load_obj_mesh(path, T.MATERIAL_DIELECTRIC)
    t = t.copy()
    t[:, 1::3] += np.float32(lift)
    tr, nr = quad(0.0, (-3.0, 3.0, -3.0, 3.0), 1.0)
    ft, fa = _assemble([(tr, nr, T.MATERIAL_DIFFUSE)])
    tris = np.concatenate([t, ft])
    att = np.concatenate([a, fa])
    lam = np.linspace(380.0, 720.0, n_lambdas).astype(np.float32)
    ext = (-1.25, 1.25, -1.25, 1.25)
    return CausticScene(tris=tris, attribs=att, p_light=(0.3, 3.0, 0.2), points=receiver_points(width, width, ext),
                        width=width, height=width, extent=ext, cauchy_a=cauchy_a, cauchy_b=cauchy_b, lambdas=lam,
                        photons_per_triangle=max(1, photons // t.shape[0]),
                        eta_min=cauchy_a + cauchy_b / 380.0 ** 2, eta_max=cauchy_a + cauchy_b / 720.0 ** 2,
                        name=f"{os.path.basename(path)} as glass, {t.shape[0]} tris, {width}^2, {photons} photons",
                        info={"path": path})


def load_obj(path: str) -> np.ndarray:
    """Triangles of a Wavefront OBJ file, (n, 9) float32, each face
    fan-triangulated in file order (v0, vi, vi+1)."""
    return load_obj_mesh(path)[0]
