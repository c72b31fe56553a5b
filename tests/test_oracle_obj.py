"""OBJ ingestion (palz_mesh_load_obj, loadFrame's rules pk.h:1246-1362): the
library's loader against the oracle restatement bit for bit (the reference's
table.obj and random meshes in every index form), and both pinned by known
answers (a quad's fan and geometric normal, negative indices, materials) and
by what the mesh fixes (triangle count sum(n - 2), fan area = polygon area by
Newell's formula)."""
import os

import numpy as np
import pytest

from oracle import obj as oobj
from palzilla_amd import scenes
from palzilla_amd import types as T

TABLE = os.path.join(os.path.dirname(__file__), "golden", "table.obj")


def _native(text, tmp_path, default_material=0, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return scenes.load_obj_mesh(str(p), default_material)


QUAD = "v 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\n"


def test_quad_fan_and_geometric_normal(tmp_path):
    for face in ("f 1 2 3 4", "f -4 -3 -2 -1", "f 1/1 2/2 3/3 4/4"):
        t, a = _native(QUAD + "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n" + face + "\n", tmp_path)
        assert t.tolist() == [[0, 0, 0, 1, 0, 0, 1, 0, 1], [0, 0, 0, 1, 0, 1, 0, 0, 1]], face
        # cross((1,0,0), (1,0,1)) = (0,-1,0), per corner
        assert np.all(a["shading_normals"] == np.array([0, -1, 0], np.float32))
        assert np.all(a["material"] == T.MATERIAL_DIFFUSE)


def test_corner_normals_and_materials(tmp_path):
    text = QUAD + "vn 0 1 0\nvn 0.6 0.8 0\nusemtl dielectric\nf 1//1 2//2 3//1\nusemtl glass\nf 1/1/2 3//1 4//2\n" \
                  "usemtl mirror\nf 2 3 4\n"
    text = text.replace("f 1/1/2", "vt 0 0\nf 1/1/2")
    t, a = _native(text, tmp_path, default_material=T.MATERIAL_DIFFUSE)
    assert a["material"].tolist() == [T.MATERIAL_DIELECTRIC, T.MATERIAL_DIFFUSE, T.MATERIAL_MIRROR]
    n = a["shading_normals"]
    assert n[0].tolist() == [[0, 1, 0], [0.6000000238418579, 0.800000011920929, 0], [0, 1, 0]]
    assert n[1].tolist() == [[0.6000000238418579, 0.800000011920929, 0], [0, 1, 0],
                             [0.6000000238418579, 0.800000011920929, 0]]
    _, a2 = _native(text, tmp_path, default_material=T.MATERIAL_DIELECTRIC)
    assert a2["material"][1] == T.MATERIAL_DIELECTRIC  # an unknown name takes the default


@pytest.mark.parametrize("bad", ["f 1 2 9", "f 1 2", "f 0 1 2", "f 1//3 2 3", "v 1 x 2", "v", "vn 1 2"])
def test_malformed_input_is_an_error(tmp_path, bad):
    with pytest.raises(Exception):
        _native(QUAD + "vn 0 1 0\n" + bad + "\n", tmp_path)
    with pytest.raises(Exception):
        oobj.load_obj(QUAD + "vn 0 1 0\n" + bad + "\n")


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(Exception):
        scenes.load_obj_mesh(str(tmp_path / "nope.obj"))


def _newell_area(P):
    n = np.zeros(3)
    for i in range(len(P)):
        a, b = P[i], P[(i + 1) % len(P)]
        n += np.cross(a, b)
    return 0.5 * np.linalg.norm(n)


def test_table_mesh_matches_the_oracle_and_its_polygons():
    text = open(TABLE).read()
    t, a = scenes.load_obj_mesh(TABLE)
    ot, oa = oobj.load_obj(text)
    assert t.tobytes() == ot.tobytes() and a.tobytes() == oa.tobytes()
    verts = np.array([[float(x) for x in l.split()[1:4]] for l in text.splitlines() if l.startswith("v ")])
    faces = [[int(c.split("/")[0]) - 1 for c in l.split()[1:]] for l in text.splitlines() if l.startswith("f ")]
    assert t.shape[0] == sum(len(f) - 2 for f in faces) == 3068
    tri = t.astype(np.float64).reshape(-1, 3, 3)
    fan = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    poly = sum(_newell_area(verts[f]) for f in faces)
    assert abs(fan - poly) <= 1e-6 * poly
    assert np.allclose(np.linalg.norm(a["shading_normals"], axis=2), 1.0, atol=1e-4)


def test_random_meshes_match_the_oracle(tmp_path):
    rng = np.random.default_rng(11)
    for trial in range(12):
        nv, nn = int(rng.integers(4, 40)), int(rng.integers(1, 10))
        lines = [f"v {x:.7g} {y:.7g} {z:.7g}" for x, y, z in rng.normal(size=(nv, 3))]
        lines += [f"vn {x:.7g} {y:.7g} {z:.7g}" for x, y, z in rng.normal(size=(nn, 3))]
        lines.append("vt 0.5 0.5")
        for _ in range(int(rng.integers(1, 30))):
            if rng.random() < 0.2:
                lines.append("usemtl " + rng.choice(["mirror", "dielectric", "paint"]))
            k = int(rng.integers(3, 8))
            corners = []
            for _ in range(k):
                vi = int(rng.integers(1, nv + 1))
                vi = vi if rng.random() < 0.7 else vi - nv - 1  # negative form
                form = rng.integers(0, 4)
                ni = int(rng.integers(1, nn + 1))
                corners.append([f"{vi}", f"{vi}/1", f"{vi}//{ni}", f"{vi}/1/{ni}"][form])
            lines.append("f " + " ".join(corners))
            if rng.random() < 0.1:
                lines.append("# comment\ns off\ng group")
        text = "\n".join(lines) + "\n"
        dm = int(rng.integers(0, 3))
        t, a = _native(text, tmp_path, dm, name=f"r{trial}.obj")
        with np.errstate(all="ignore"):
            ot, oa = oobj.load_obj(text, dm)
        assert t.tobytes() == ot.tobytes(), trial
        assert a.tobytes() == oa.tobytes(), trial
