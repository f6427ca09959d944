"""Scene ingest (Scene::Load restatement, rt.hpp:1519-2154) and generators."""
import os

import numpy as np
import pytest
import yaml

import nanogi_amd as ngi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CORNELL = os.path.join(GOLDEN, "cornelbox", "scene.yml")
RUNC = os.path.join(GOLDEN, "runc_data", "scene.yml")


def _obj_triangles(path):
    """Fan-triangulated face count of the first object in an OBJ file."""
    n = 0
    seen_faces = False
    for line in open(path):
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "o" and seen_faces:
            break
        if tok[0] == "f":
            n += len(tok) - 3
            seen_faces = True
    return n


@pytest.mark.parametrize("path", [CORNELL, RUNC])
def test_yaml_fixture_matches_independent_parse(path):
    s = ngi.Scene.load_yaml(path, 1.5)
    d = s.desc()
    ref = yaml.safe_load(open(path))
    prims = ref["scene"]["primitives"]
    assert len(d["primitives"]) == len(prims)
    letters = {"D": ngi.PRIM_D, "G": ngi.PRIM_G, "S": ngi.PRIM_S, "L": ngi.PRIM_L, "E": ngi.PRIM_E}
    total = 0
    for got, want in zip(d["primitives"], prims):
        assert got["type"] == sum(letters[c] for c in want["type"])
        if "mesh" in want:
            n = _obj_triangles(os.path.join(os.path.dirname(path), want["mesh"]["path"]))
            assert got["tri_count"] == n
            assert got["tri_begin"] == total
            total += n
        else:
            assert got["tri_count"] == 0
        par = want["params"]
        if "L" in par:
            np.testing.assert_array_equal(got["Le"], par["L"]["area"]["Le"])
        if "D" in par:
            np.testing.assert_array_equal(got["D_R"], par["D"]["R"])
        if "G" in par:
            np.testing.assert_array_equal(got["G_R"], par["G"]["R"])
            np.testing.assert_array_equal(got["G_eta"], par["G"]["Eta"])
            np.testing.assert_array_equal(got["G_k"], par["G"]["K"])
            assert got["G_roughness"] == par["G"]["Roughness"]
        if "S" in par:
            f = par["S"][par["S"]["type"]]
            np.testing.assert_array_equal(got["S_R"], f["R"])
            assert (got["S_eta1"], got["S_eta2"]) == (f["eta1"], f["eta2"])
        if "E" in par:
            ph = par["E"]["pinhole"]
            np.testing.assert_array_equal(got["E_eye"], ph["view"]["eye"])
            np.testing.assert_array_equal(got["E_center"], ph["view"]["center"])
            assert got["E_fov_deg"] == ph["perspective"]["fov"]
    assert d["faces"].shape[0] == total
    assert d["aspect"] == 1.5


def test_cornell_fixture_has_38_triangles():
    assert ngi.Scene.load_yaml(CORNELL, 1.0).num_triangles == 38


def test_obj_vertices_and_normals_are_kept():
    s = ngi.Scene.load_yaml(CORNELL, 1.0)
    d = s.desc()
    light = d["primitives"][0]
    tri = d["faces"][light["tri_begin"]]
    np.testing.assert_allclose(d["normals"][tri], [[0, -1, 0]] * 3)
    np.testing.assert_allclose(d["positions"][tri[0]], [343.0, 548.749927, 227.000092])


def _write(tmp_path, text, files=()):
    for name, content in files:
        (tmp_path / name).write_text(content)
    p = tmp_path / "scene.yml"
    p.write_text(text)
    return str(p)


QUAD = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n"
CAM = """    - type: [E]
      params:
        E:
          type: pinhole
          pinhole:
            We: [1, 1, 1]
            view: {eye: [0, 0, 5], center: [0, 0, 0], up: [0, 1, 0]}
            perspective:
              fov: 45
"""


@pytest.mark.parametrize("text,status", [
    ("version: 2\nscene:\n  primitives:\n" + CAM, ngi.ERR_VERSION),
    ("version: 6\nscene:\n  primitives:\n" + CAM, ngi.ERR_VERSION),
    ("version: 5\nscene:\n  primitives:\n    - type: [L, E]\n      params: {}\n", ngi.ERR_INVALID_TYPE),
    ("version: 5\nscene:\n  primitives:\n    - type: [X]\n      params: {}\n", ngi.ERR_INVALID_TYPE),
    ("version: 5\nscene:\n  primitives:\n    - type: [L]\n      params:\n        L:\n          type: area\n"
     "          area:\n            Le: [1, 1, 1]\n", ngi.ERR_INVALID_SCENE),
    ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'missing.obj'\n"
     "      params:\n        D:\n          R: [1, 1, 1]\n", ngi.ERR_IO),
    ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'q.obj'\n"
     "      params:\n        D:\n          TexR: 'tex.png'\n", ngi.ERR_UNSUPPORTED),
    ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'q.obj'\n"
     "      params:\n        D: {}\n", ngi.ERR_INVALID_SCENE),
    ("version: 5\nscene:\n  primitives:\n    - type: [D\n", ngi.ERR_PARSE),
    ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'q.obj'\n"
     "      params:\n        D:\n          R: [1, one, 1]\n", ngi.ERR_PARSE),
])
def test_load_errors(tmp_path, text, status):
    path = _write(tmp_path, text, [("q.obj", QUAD)])
    with pytest.raises(ngi.NgiError) as e:
        ngi.Scene.load_yaml(path, 1.0)
    assert e.value.status == status


def test_missing_scene_file(tmp_path):
    with pytest.raises(ngi.NgiError) as e:
        ngi.Scene.load_yaml(str(tmp_path / "nope.yml"), 1.0)
    assert e.value.status == ngi.ERR_IO


def test_quad_is_fan_triangulated_and_flat_normals_generated(tmp_path):
    text = ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'q.obj'\n"
            "      params:\n        D:\n          R: [1, 1, 1]\n" + CAM)
    s = ngi.Scene.load_yaml(_write(tmp_path, text, [("q.obj", QUAD)]), 1.0)
    d = s.desc()
    assert d["faces"].shape == (2, 3)
    P = d["positions"][d["faces"]]
    np.testing.assert_allclose(P[0], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    np.testing.assert_allclose(P[1], [[0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_allclose(d["normals"][d["faces"]].reshape(-1, 3), [[0, 0, 1]] * 6)


def test_only_first_object_is_loaded(tmp_path):
    obj = QUAD + "o second\nv 5 5 5\nv 6 5 5\nv 6 6 5\nf 5 6 7\n"
    text = ("version: 5\nscene:\n  primitives:\n    - type: [D]\n      mesh:\n        path: 'q.obj'\n"
            "      params:\n        D:\n          R: [1, 1, 1]\n" + CAM)
    s = ngi.Scene.load_yaml(_write(tmp_path, text, [("q.obj", obj)]), 1.0)
    assert s.num_triangles == 2


@pytest.mark.parametrize("kind,detail,expect", [
    (ngi.GEN_CORNELL, 0, 36),
    (ngi.GEN_NAVE, 0, 76842),
    (ngi.GEN_SPHERE_FIELD, 3, 20 * 4 ** 3 + 2 * 63 ** 2 + 320 + 32),
    (ngi.GEN_SPHERE_FIELD, 8, 20 * 4 ** 8 + 2 * 511 ** 2 + 320 + 32),
])
def test_generator_triangle_counts(kind, detail, expect):
    if kind == ngi.GEN_SPHERE_FIELD and detail == 8:
        pass  # 1.8M triangles; generation takes well under a second
    assert ngi.Scene.generate(kind, 1, 1.0, detail).num_triangles == expect


@pytest.mark.parametrize("kind", [ngi.GEN_CORNELL, ngi.GEN_NAVE])
def test_generated_normals_agree_with_winding(kind):
    """Vertex normals lie on the winding side of every triangle (one-sided
    surfaces, rt.hpp:994-1005, would otherwise go black)."""
    d = ngi.Scene.generate(kind, 1, 1.0).desc()
    P, N, F = d["positions"], d["normals"], d["faces"]
    gn = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    for k in range(3):
        assert np.all(np.einsum("ij,ij->i", gn, N[F[:, k]]) > 0)


def test_generated_positions_are_float_exact():
    d = ngi.Scene.generate(ngi.GEN_NAVE, 1, 1.0).desc()
    np.testing.assert_array_equal(d["positions"], d["positions"].astype(np.float32).astype(np.float64))


def test_nave_materials_and_lights():
    d = ngi.Scene.generate(ngi.GEN_NAVE, 5, 1.0).desc()
    pr = d["primitives"]
    diffuse = pr[pr["type"] == ngi.PRIM_D]
    lights = pr[pr["type"] == ngi.PRIM_L]
    assert len(diffuse) == 6 and len(lights) == 4
    assert np.all((diffuse["D_R"] >= 0.2) & (diffuse["D_R"] <= 0.8))
    np.testing.assert_array_equal(lights["Le"], np.full((4, 3), 20.0))


def test_sphere_field_roughness_mapping():
    d = ngi.Scene.generate(ngi.GEN_SPHERE_FIELD, 2, 1.0, 4).desc()
    g = d["primitives"][d["primitives"]["type"] == ngi.PRIM_G]
    assert len(g) == 20 * 4 ** 4 // 64
    n = 2.0 / g["G_roughness"] ** 2 - 2.0   # alpha = sqrt(2 / (n + 2))
    assert n.min() >= 10 - 1e-6 and n.max() <= 1000 + 1e-6
    assert np.all(g["tri_count"] == 64)


def test_save_and_reload_round_trip(tmp_path):
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    s.save_yaml(str(tmp_path / "out"))
    r = ngi.Scene.load_yaml(str(tmp_path / "out" / "scene.yml"), 1.0)
    a, b = s.desc(), r.desc()
    np.testing.assert_allclose(a["positions"][a["faces"]], b["positions"][b["faces"]])
    np.testing.assert_allclose(a["normals"][a["faces"]], b["normals"][b["faces"]])
    for f in ("type", "tri_begin", "tri_count", "Le", "D_R", "E_eye", "E_center", "E_fov_deg"):
        np.testing.assert_array_equal(a["primitives"][f], b["primitives"][f])


def test_set_params():
    s = ngi.Scene.generate(ngi.GEN_CORNELL, 1, 1.0)
    n = len(s.desc()["primitives"])
    refl = np.full((n, 3), 0.25)
    Le = np.full((n, 3), 3.0)
    s.set_params(refl, Le)
    pr = s.desc()["primitives"]
    assert np.all(pr["D_R"][pr["type"] == ngi.PRIM_D] == 0.25)
    assert np.all(pr["Le"][pr["type"] == ngi.PRIM_L] == 3.0)
    assert np.all(pr["Le"][pr["type"] == ngi.PRIM_D] == 0.0)


def test_create_validates_arrays():
    from tests import scenes
    B = scenes.SceneBuilder()
    p = B.prim(type=ngi.PRIM_D, D_R=(1, 1, 1))
    B.rect(p, (0, 0, 0), (0, 1, 0), (1, 0, 0), 1, 1)
    B.faces[0][0] = 99
    with pytest.raises(ngi.NgiError) as e:
        B.scene()
    assert e.value.status == ngi.ERR_INVALID_ARG
