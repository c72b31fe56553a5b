"""Oracle: Wavefront OBJ ingestion with the rules of the reference's
loadFrame (pk.h:1246-1362) — TEST INFRASTRUCTURE, a pure-Python restatement
of the library's loader, pinned by the tests against known answers and
against what the mesh fixes (triangle counts, polygon areas).

Per face corner: the position and the shading normal; the material of the
`usemtl` in force ("mirror" 1, "dielectric" 2, otherwise the default); faces
of n corners become the fan (0, i, i + 1); a corner without a normal takes
normalize(cross(v1 - v0, v2 - v0)) of its triangle in float32; numbers are
parsed as double and rounded to float32; indices are 1-based or negative
(relative to the end of the list)."""
from __future__ import annotations

import numpy as np

ATTRIB = np.dtype([("material", "<i4"), ("shading_normals", "<f4", (3, 3))])  # TriangleAttrib (pk.h:74-78)


def _index(tok: str, n: int) -> int:
    i = int(tok)
    if i == 0:
        raise ValueError("index 0")
    r = i - 1 if i > 0 else n + i
    if not 0 <= r < n:
        raise ValueError("index out of range")
    return r


def load_obj(text: str, default_material: int = 0):
    """(triangles (n, 9) float32, attribs) of OBJ text."""
    f32 = np.float32
    v, vn, tris, att = [], [], [], []
    material = default_material
    for raw in text.splitlines():
        parts = raw.split()
        if not parts:
            continue
        key = parts[0]
        if key in ("v", "vn"):
            if len(parts) < 4:
                raise ValueError(f"{key} needs three numbers")
            (v if key == "v" else vn).append([f32(float(x)) for x in parts[1:4]])
        elif key == "usemtl":
            name = parts[1] if len(parts) > 1 else ""
            material = {"mirror": 1, "dielectric": 2}.get(name, default_material)
        elif key == "f":
            corners = []
            for c in parts[1:]:
                fields = c.split("/")
                p = _index(fields[0], len(v))
                nrm = _index(fields[2], len(vn)) if len(fields) == 3 and fields[2] else None
                corners.append((p, nrm))
            if len(corners) < 3:
                raise ValueError("face with fewer than 3 corners")
            for j in range(1, len(corners) - 1):
                tri = [corners[0], corners[j], corners[j + 1]]
                P = [np.array(v[p], dtype=f32) for p, _ in tri]
                ng = None
                if any(n is None for _, n in tri):
                    e1, e2 = P[1] - P[0], P[2] - P[0]
                    g = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]], dtype=f32)
                    ln = np.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
                    ng = np.array([g[0] / ln, g[1] / ln, g[2] / ln], dtype=f32)
                tris.append(np.concatenate(P))
                att.append(([np.array(vn[n], dtype=f32) if n is not None else ng for _, n in tri], material))
    t = np.asarray(tris, dtype=f32).reshape(-1, 9)
    a = np.zeros(len(att), dtype=ATTRIB)
    for i, (ns, m) in enumerate(att):
        a["shading_normals"][i] = np.asarray(ns, dtype=f32)
        a["material"][i] = m
    return t, a
