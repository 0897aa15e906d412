"""The reference's own scene files load through the drop-in front door.

Every Scene/**/*.json of the reference tree (skipped where the tree is absent)
is copied to a temporary directory byte for byte, except that each asset path
(OBJ models, textures, environment maps: the reference ships none of them) is
replaced by the path of a generated stand-in.  Each copy must load, and what
it declares must come out as the reference's Scene::ParseSceneFile
(Src/Core/scene.cpp:129-644) builds it: camera values, every material with the
reference's per-type defaults, the object -> material bindings (material,
material_overrides, the OBJ's usemtl name, default_material;
scene.cpp:12-36), lights and textures.  The library's loader is also compared
with the oracle's independent parser on the same files.  No copy of a
reference scene is kept."""
import glob
import json
import os

import numpy as np
import pytest

import cudaraytracer_amd as rt
from cudaraytracer_amd import scenes

REF_SCENES = sorted(glob.glob("/root/reference/Scene/**/*.json", recursive=True))
pytestmark = pytest.mark.skipif(not REF_SCENES, reason="the reference tree is not present")

TYPE = {"light": 0, "diffuse": 1, "pbr": 2, "dielectric": 3, "glass": 3, "subsurface": 4}


def _cube_obj(path, groups):
    """One unit cube per (object name, usemtl name or None), with normals and
    texture coordinates, each cube shifted along x."""
    lines = []
    corners = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)])
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    for g, (oname, mtl) in enumerate(groups):
        base = 8 * g
        lines.append(f"o {oname}")
        for c in corners:
            lines.append("v %g %g %g" % (c[0] + 1.5 * g, c[1], c[2]))
        if mtl is not None:
            lines.append(f"usemtl {mtl}")
        for a, b, c, d in faces:
            lines.append(f"f {base + a + 1} {base + b + 1} {base + c + 1}")
            lines.append(f"f {base + a + 1} {base + c + 1} {base + d + 1}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _materialise(ref_json, tmp):
    """Copy a reference scene with its asset paths pointed at stand-ins;
    returns (copy path, parsed document, stand-in usemtl names per object)."""
    raw = open(ref_json, "rb").read()
    doc = json.loads(raw.decode("utf-8-sig"))
    text = raw
    usemtl = {}
    mat_names = [m["name"] for m in doc["materials"] if "name" in m]
    made = {}

    def replace(old, new):
        nonlocal text
        text = text.replace(json.dumps(old).encode(), json.dumps(new).encode())

    for i, t in enumerate(doc.get("textures", [])):
        if "path" in t and t["path"] not in made:
            p = os.path.join(tmp, f"tex{i}.png")
            rng = np.random.default_rng(i)
            scenes.write_png_rgba(p, rng.integers(0, 256, (4, 4, 4)).astype(np.uint8))
            made[t["path"]] = p
            replace(t["path"], p)
    for i, o in enumerate(doc["objects"]):
        if "path" not in o:
            continue
        # one group per object, or (for objects that bind through the OBJ's material
        # names) a group naming a scene material and a group naming none
        if "material" in o or "material_overrides" not in o and "default_material" not in o:
            groups = [("part", None)]
        else:
            groups = [("a", mat_names[0]), ("b", "no_such_material")]
        usemtl[o["name"]] = groups
        p = os.path.join(tmp, f"obj{i}.obj")
        _cube_obj(p, groups)
        # every object gets its own stand-in, so paths are replaced one occurrence at a time
        old = json.dumps(o["path"]).encode()
        k = text.index(old)
        text = text[:k] + json.dumps(p).encode() + text[k + len(old):]
    for i, L in enumerate(doc["lights"]):
        if L.get("type") == "env" and "path" in L and L["path"] not in made:
            p = os.path.join(tmp, f"env{i}.hdr")
            scenes.write_hdr(p, scenes.sky_env(64, 32))
            made[L["path"]] = p
            replace(L["path"], p)
    out = os.path.join(tmp, os.path.basename(ref_json))
    with open(out, "wb") as f:
        f.write(text)
    return out, doc, usemtl


def _expected_material(m, tex_names):
    """(type, basecolor, emission, roughness, metallic, transmission, ior,
    scatter, texture) as Scene::ParseMaterials sets them (scene.cpp:254-395)
    over the Material member defaults (material.h:57-71)."""
    t = TYPE[m["type"]]
    bc, ke, rough, metal, trans, ior, sc = [0, 0, 0], [0, 0, 0], 0.5, 0.0, 1.0, 1.0, [0, 0, 0]
    if t == 1:
        bc = m.get("basecolor", [0, 0, 0])
    elif t == 0:
        ke = m.get("emission", [0, 0, 0])
    elif t == 2:
        bc = m.get("basecolor", [0, 0, 0])
        metal = m.get("metallic", 0.0)
        rough = m.get("roughness", 0.5)
    elif t == 4:
        bc = m.get("basecolor", [0, 0, 0])
        sc = m.get("scatter_distance", [1, 1, 1])
        ior = m.get("ior", ior)
        rough = m.get("roughness", rough)
        metal = 0.0
    elif t == 3:
        bc = m.get("basecolor", [1, 1, 1])
        ior = m.get("ior", 1.5)
        trans = m.get("transmission", 1.0)
        rough = m.get("roughness", 0.0)
        metal = 0.0
    tex = tex_names.get(m.get("diffuse_texture"), -1)
    return t, bc, ke, rough, metal, trans, ior, sc, tex


def test_all_reference_scenes_are_found():
    assert len(REF_SCENES) == 25
    assert any(open(p, "rb").read(3) == b"\xef\xbb\xbf" for p in REF_SCENES)  # one starts with a BOM


@pytest.mark.parametrize("ref_json", REF_SCENES, ids=[os.path.relpath(p, "/root/reference/Scene") for p in REF_SCENES])
def test_reference_scene_loads_with_reference_semantics(tmp_path, oracle_lib, ref_json):
    path, doc, usemtl = _materialise(ref_json, str(tmp_path))
    scene = rt.Scene(path, bvh_builder=rt.BVH_HOST)
    info = scene.info()
    cam = doc["camera"]
    # camera (ParseCamera, scene.cpp:187-252); "target" and "distance" are not parsed
    assert (info.width, info.height) == tuple(cam["resolution"])
    assert info.spp == cam["spp"] and info.max_bounces == cam["max_bounces"]
    assert info.rr == np.float32(cam["rr"])
    assert info.denoise == int(bool(cam.get("denoise", False)))
    # materials, with every per-type default
    tex_names = {t["name"]: i for i, t in enumerate(doc["textures"]) if "name" in t and "path" in t}
    mats = scene.materials()
    assert mats.shape[0] == len(doc["materials"]) == info.num_materials
    ints = mats.view(np.int32)
    for i, m in enumerate(doc["materials"]):
        t, bc, ke, rough, metal, trans, ior, sc, tex = _expected_material(m, tex_names)
        assert ints[i, 0] == t and ints[i, 15] == 1, m["name"]
        assert np.array_equal(mats[i, 1:4], np.float32(bc)) and np.array_equal(mats[i, 4:7], np.float32(ke))
        assert mats[i, 7] == np.float32(rough) and mats[i, 8] == np.float32(metal), m["name"]
        assert mats[i, 9] == np.float32(trans) and mats[i, 10] == np.float32(ior), m["name"]
        assert np.array_equal(mats[i, 11:14], np.float32(sc)) and ints[i, 14] == tex, m["name"]
    # object -> material bindings (ResolveMaterialName, scene.cpp:12-36)
    index = {m["name"]: i for i, m in enumerate(doc["materials"])}
    mesh4, _ = scene.meshes()
    expected = []
    for o in doc["objects"]:
        for _, mtl in usemtl[o["name"]]:
            if "material" in o:
                name = o["material"]
            elif isinstance(o.get("material_overrides"), dict) and mtl in o["material_overrides"]:
                name = o["material_overrides"][mtl]
            elif mtl in index:
                name = mtl
            else:
                name = o.get("default_material", "")
            if name in index:
                expected.append(index[name])
    assert mesh4[:, 0].tolist() == expected
    # lights: area lights own their mesh, one env light decodes its map
    assert info.num_lights == len(doc["lights"])
    for i, L in enumerate(doc["lights"]):
        if L["type"] == "area":
            assert (mesh4[:, 1] == i).sum() == 1
    has_env = any(L["type"] == "env" for L in doc["lights"])
    assert scene.envmap().shape == ((32, 64, 3) if has_env else (0, 0, 3))
    assert info.num_textures == len(doc["textures"])
    # the oracle's own parser reads the same scene to the same arrays
    a = oracle_lib.parse_scene(path)
    pos = np.zeros((info.num_triangles, 9), np.float32)
    import ctypes

    rt.lib().rt_scene_get_triangles(scene._h, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), None, None, None)
    assert np.array_equal(pos.view(np.int32), a.pos.view(np.int32))
    assert mesh4[:, 0].tolist() == a.mesh_material[: len(mesh4)].tolist()
    assert np.array_equal(ints[:, 0], a.materials[: len(ints), 0].astype(np.int32))
    assert np.array_equal(mats[:, 1:14], a.materials[: len(mats), 1:14])
