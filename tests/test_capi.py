"""The C-ABI library loads, exports every symbol include/rt.h declares, and
reports errors the way the reference's loader does (fatal scene errors)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import cudaraytracer_amd as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "rt.h")).read()
    return sorted(set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", text)))


def test_header_and_library_agree():
    lib = ctypes.CDLL(rt.LIB_PATH)
    declared = _declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(rt.EXPORTS) == declared


def test_load_errors_are_reported(tmp_path):
    with pytest.raises(rt.RtError) as e:
        rt.Scene(str(tmp_path / "missing.json"))
    assert e.value.code == -2  # RT_ERR_IO
    bad = tmp_path / "bad.json"
    bad.write_text("{ not json")
    with pytest.raises(rt.RtError) as e:
        rt.Scene(str(bad))
    assert e.value.code == -3  # RT_ERR_SCENE
    nocam = tmp_path / "nocam.json"
    nocam.write_text(json.dumps({"camera": {}, "textures": [], "materials": [], "objects": [], "lights": []}))
    with pytest.raises(rt.RtError, match="type"):
        rt.Scene(str(nocam))


def test_missing_light_mesh_and_empty_scene(tmp_path, scene_dir):
    doc = json.load(open(scene_dir["cornell_small"]))
    base = os.path.dirname(scene_dir["cornell_small"])
    for o in doc["objects"]:
        o["path"] = os.path.join(base, o["path"])
    doc["lights"][0]["lightMeshName"] = "nope"
    p = tmp_path / "l.json"
    p.write_text(json.dumps(doc))
    with pytest.raises(rt.RtError, match="light mesh"):
        rt.Scene(str(p))
    doc["objects"] = []
    doc["lights"] = []
    p.write_text(json.dumps(doc))
    with pytest.raises(rt.RtError, match="no triangles"):
        rt.Scene(str(p))


def test_scene_info_and_material_defaults(tmp_path, scene_dir):
    s = rt.Scene(scene_dir["cornell_spheres"])
    info = s.info()
    assert (info.width, info.height, info.spp, info.max_bounces) == (512, 512, 64, 12)
    assert info.num_triangles == 36 + 4 * 5120
    assert info.seed == 1234 and info.rr_start_depth == 3
    mesh4, area = s.meshes()
    light_mesh = [i for i in range(info.num_meshes) if mesh4[i, 1] >= 0]
    assert len(light_mesh) == 1 and area[light_mesh[0]] == pytest.approx(0.25, rel=1e-6)


def test_transform_and_rotation_match_the_oracle_parse(tmp_path, oracle_lib):
    """Translate / rotate / scale (T R S, scene.cpp:421-448) and normals through
    the inverse transpose (objloader.cpp:36-41): library vs the oracle's parse."""
    from cudaraytracer_amd import scenes

    V, F = scenes.icosphere(1)
    spec = scenes.SceneSpec(camera=scenes._camera(16, 16, 1, 2), materials=scenes._base_materials(),
                            meshes=[scenes.MeshSpec("ball", "white", V[F].astype(np.float32),
                                                    V[F].astype(np.float32))], lights=[])
    path = scenes.write_scene(spec, str(tmp_path), "t")
    doc = json.load(open(path))
    doc["objects"][0].update({"translate": [0.5, -1, 2], "scale": [2, 2, 2], "rotate": [{"rotateY": 30}]})
    json.dump(doc, open(path, "w"))
    s = rt.Scene(path)
    pos, nrm, uv, mesh = s.triangles()
    a = oracle_lib.parse_scene(path)
    assert np.array_equal(pos, a.pos)
    assert np.allclose(nrm, a.nrm, atol=2e-7)
    c = pos.reshape(-1, 3).mean(axis=0)
    assert np.allclose(c, [0.5, -1, 2], atol=1e-5)
