"""Diffuse textures on the GPU against the oracle: the lookup itself
(Texture::Sample, Src/Core/texture.cu:242-249) at edge and random coordinates,
and whole textured renders, bit for bit; the dielectric ignores its texture
(material.cu:190) in the product as in the oracle."""
import json
import os
import shutil

import numpy as np
import pytest

import cudaraytracer_amd as rt

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def test_texture_lookup_bit_exact(scene_dir, oracle_lib):
    path = scene_dir["textured_cornell"]
    scene = rt.Scene(path)
    rs = oracle_lib.RefScene(oracle_lib.parse_scene(path))
    rng = np.random.default_rng(7)
    edges = np.array([0.0, -0.0, 1.0, 0.999999, 1e-7, -1e-7, 0.5, -3.0, 4.0, np.nan, np.inf, -np.inf], np.float32)
    grid = np.stack(np.meshgrid(edges, edges), -1).reshape(-1, 2)
    uv = np.concatenate([grid, rng.uniform(-0.25, 1.25, (4000, 2)).astype(np.float32)])
    for tex in range(scene.info().num_textures):
        print(f"texture {tex}: {scene.texture(tex).shape}", flush=True)
        got = scene.texture_lookup(tex, uv)
        want = np.array([rs.texture_albedo(tex, float(u), float(v)) for u, v in uv], np.float32)
        assert np.array_equal(_bits(got), _bits(want)), f"texture {tex}: {np.count_nonzero(got != want)} differ"


def _render(path, spp):
    r = rt.Renderer(rt.Scene(path))
    r.render(spp)
    return (r.read_accum(rt.RT_ACCUM_DIRECT).reshape(-1, 4), r.read_accum(rt.RT_ACCUM_INDIRECT).reshape(-1, 4),
            r.stats())


def test_textured_render_bit_exact(scene_dir, oracle_lib):
    path = scene_dir["textured_cornell"]
    d, i, st = _render(path, 4)
    rs = oracle_lib.RefScene(oracle_lib.parse_scene(path))
    rd, ri, rst = rs.render(4, rs.params())
    assert np.array_equal(_bits(d[:, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[:, :3]), _bits(ri[:, :3]))
    assert st.extension_rays == rst[0] and st.shadow_rays == rst[1]


def _variant(src_json, tmp, drop):
    """Copy of a scene directory with the diffuse_texture key removed from the named materials."""
    src_dir = os.path.dirname(src_json)
    dst_dir = os.path.join(tmp, "variant")
    shutil.copytree(src_dir, dst_dir)
    dst = os.path.join(dst_dir, os.path.basename(src_json))
    with open(dst) as f:
        doc = json.load(f)
    for m in doc["materials"]:
        if m["name"] in drop:
            m.pop("diffuse_texture", None)
    with open(dst, "w") as f:
        json.dump(doc, f)
    return dst


def test_dielectric_texture_has_no_effect(scene_dir, tmp_path):
    path = scene_dir["textured_cornell"]
    d, i, _ = _render(path, 2)
    d2, i2, _ = _render(_variant(path, str(tmp_path / "a"), {"glass_tex"}), 2)
    assert np.array_equal(_bits(d), _bits(d2)) and np.array_equal(_bits(i), _bits(i2))
    d3, _, _ = _render(_variant(path, str(tmp_path / "b"), {"wall_tex"}), 2)
    assert not np.array_equal(_bits(d), _bits(d3))
