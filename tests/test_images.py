"""Texture images: the library's decoders and the oracle's reader against the
texels the reference's own loader (stb_image 2.30, as Texture::Load calls it,
Src/Core/texture.cu:207-214) returns for the same files, recorded under
tests/golden/images/ by make_image_fixtures.py; loader error paths; and the
oracle's texture lookup (Texture::Sample, texture.cu:242-249) against what its
addressing fixes."""
import json
import os
import struct

import numpy as np
import pytest

import cudaraytracer_amd as rt
from oracle import rtref
from oracle.rtref_image import read_texture

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "images")
EXPECTED = np.load(os.path.join(GOLDEN, "expected.npz"))
NAMES = sorted(EXPECTED.files)


def _scene_with_textures(tmp, textures, materials=None):
    """A one-quad scene whose "textures" array is `textures` (name, path) pairs."""
    with open(os.path.join(tmp, "quad.obj"), "w") as f:
        f.write("v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                "f 1/1 3/3 2/2\nf 1/1 4/4 3/3\n")
    doc = {
        "camera": {"type": "pinhole", "resolution": [8, 8], "fov": 40, "spp": 1, "max_bounces": 2, "rr": 0.8,
                   "direction": [0, -1, 0]},
        "textures": [{"name": n, "path": p} for n, p in textures],
        "materials": materials or [{"name": "m", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5]}],
        "objects": [{"name": "quad", "path": "quad.obj", "material": (materials or [{"name": "m"}])[0]["name"]}],
        "lights": [],
    }
    path = os.path.join(tmp, "scene.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    return path


def test_fixtures_cover_the_formats():
    kinds = {n.split("_")[0] for n in NAMES}
    assert kinds == {"png", "jpg", "pnm"}
    assert len(NAMES) >= 50
    assert all(EXPECTED[n].dtype == np.uint8 and EXPECTED[n].shape[2] == 4 for n in NAMES)


def test_library_decoders_match_reference_loader(tmp_path):
    path = _scene_with_textures(str(tmp_path), [(n, os.path.join(GOLDEN, n)) for n in NAMES])
    scene = rt.Scene(path)
    assert scene.info().num_textures == len(NAMES)
    for i, n in enumerate(NAMES):
        got = scene.texture(i)
        assert got.shape == EXPECTED[n].shape, n
        assert np.array_equal(got, EXPECTED[n]), f"{n}: {np.count_nonzero(got != EXPECTED[n])} bytes differ"


@pytest.mark.parametrize("name", [n for n in NAMES if not n.startswith("jpg")])
def test_oracle_reader_matches_reference_loader(name):
    assert np.array_equal(read_texture(os.path.join(GOLDEN, name)), EXPECTED[name])


def _bad_files(tmp):
    png = open(os.path.join(GOLDEN, "png_rgb8.png"), "rb").read()
    jpg = open(os.path.join(GOLDEN, "jpg_420_q50.jpg"), "rb").read()
    sof = jpg.index(b"\xff\xc0")
    files = {
        "truncated.png": (png[:len(png) // 2], "PNG"),
        "bad_signature.png": (b"\x89PNX" + png[4:], "unsupported image format"),
        "short_data.png": (png[:33] + png[-12:], "PNG: no image data"),
        "gif.gif": (b"GIF89a" + bytes(32), "unsupported image format"),
        "soi_only.jpg": (b"\xff\xd8\xff\xd9", "JPEG"),
        "arithmetic.jpg": (jpg[:sof] + b"\xff\xc9" + jpg[sof + 2:], "unsupported coding process"),
        "twelve_bit.jpg": (jpg[:sof + 4] + b"\x0c" + jpg[sof + 5:], "only 8-bit"),
        "empty.ppm": (b"P6\n0 4\n255\n", "zero-sized"),
    }
    out = {}
    for n, (data, msg) in files.items():
        p = os.path.join(tmp, n)
        with open(p, "wb") as f:
            f.write(data)
        out[p] = msg
    out[os.path.join(tmp, "does_not_exist.png")] = "cannot open image"
    return out


def test_unreadable_textures_fail_the_load_with_the_file_named(tmp_path):
    for p, msg in _bad_files(str(tmp_path)).items():
        scene_json = _scene_with_textures(str(tmp_path), [("t", p)])
        with pytest.raises(rt.RtError) as e:
            rt.Scene(scene_json)
        text = str(e.value)
        assert msg in text and os.path.basename(p) in text, text


def test_texture_slots_and_missing_names(tmp_path):
    """An entry without 'path' keeps its slot empty; a material naming an
    unknown texture keeps its base colour (scene.cpp:380-387, 630-633)."""
    good = os.path.join(GOLDEN, "png_rgba8.png")
    mats = [{"name": "m", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5], "diffuse_texture": "nope"}]
    path = _scene_with_textures(str(tmp_path), [("a", good)], materials=mats)
    with open(path) as f:
        doc = json.load(f)
    doc["textures"].insert(0, {"name": "no_path"})
    with open(path, "w") as f:
        json.dump(doc, f)
    scene = rt.Scene(path)
    assert scene.info().num_textures == 2
    assert scene.texture(0).shape == (0, 0, 4)
    assert np.array_equal(scene.texture(1), EXPECTED["png_rgba8.png"])
    a = rtref.parse_scene(path)
    assert a.materials[0, 14] == 0  # no texture
    assert len(a.textures) == 2 and a.textures[0].size == 0


# ------------------------------------------------------------ texture lookup
def _lookup_scene(tmp, img):
    from cudaraytracer_amd import scenes

    p = os.path.join(tmp, "t.png")
    scenes.write_png_rgba(p, img)
    mats = [{"name": "m", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5], "diffuse_texture": "t"}]
    return _scene_with_textures(tmp, [("t", p)], materials=mats)


def _ref(tmp, img):
    return rtref.RefScene(rtref.parse_scene(_lookup_scene(tmp, img)))


def test_lookup_texel_centres_are_exact(tmp_path, oracle_lib):
    """Power-of-two sizes make texel-centre coordinates exact, so the filter
    must return the texel itself: this fixes the half-texel offset, the V flip
    (v = 1 is the image's top row) and the row/column order."""
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    rs = _ref(str(tmp_path), img)
    for j in range(4):
        for i in range(8):
            u, v = (i + 0.5) / 8, 1.0 - (j + 0.5) / 4
            got = rs.texture_albedo(0, u, v)
            assert np.array_equal(got, img[j, i, :3].astype(np.float32) / np.float32(255)), (i, j)


def test_lookup_filter_wrap_and_clamp(tmp_path, oracle_lib):
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    rs = _ref(str(tmp_path), img)
    t = img[..., :3].astype(np.float64) / 255
    # Halfway between two texel centres in u: the mean of the pair.
    for j, i in ((0, 0), (2, 5), (3, 6)):
        got = rs.texture_albedo(0, (i + 1) / 8, 1.0 - (j + 0.5) / 4)
        assert np.allclose(got, (t[j, i] + t[j, i + 1]) / 2, atol=1e-6)
    # u = 0 lies halfway between the last and the first column (U wraps).
    got = rs.texture_albedo(0, 0.0, 1.0 - 2.5 / 4)
    assert np.allclose(got, (t[2, 7] + t[2, 0]) / 2, atol=1e-6)
    # Beyond the top and bottom rows V clamps: v >= 1 reads row 0, v <= 0 the last row.
    for v, row in ((1.0, 0), (7.5, 0), (0.0, 3), (-2.0, 3)):
        got = rs.texture_albedo(0, 2.5 / 8, v)
        assert np.allclose(got, t[row, 2], atol=3e-6), (v, row)
    # u is clamped to [0, 0.999999] before wrapping; NaN reads as 0.
    for u, same in ((-3.0, 0.0), (5.0, 0.999999), (float("nan"), 0.0), (float("inf"), 0.999999)):
        assert np.array_equal(rs.texture_albedo(0, u, 0.3), rs.texture_albedo(0, same, 0.3))
    # A constant texture is returned everywhere (to the rounding of two FP32 lerps).
    const = np.zeros((3, 5, 4), np.uint8)
    const[..., :3] = (17, 200, 255)
    rc = _ref(str(tmp_path), const)
    for u, v in rng.uniform(-0.5, 1.5, (50, 2)):
        assert np.allclose(rc.texture_albedo(0, u, v), np.array([17, 200, 255]) / 255, rtol=0, atol=3e-7)


def test_dielectric_ignores_its_texture_in_the_oracle(scene_dir, oracle_lib):
    """material.cu:190: the dielectric's weight reads the base colour, never the texture."""
    a = oracle_lib.parse_scene(scene_dir["textured_cornell"])
    rs = oracle_lib.RefScene(a)
    d, i, _ = rs.render(1, rs.params())
    glass = [k for k, m in enumerate(a.materials) if m[0] == 3]
    assert glass and all(a.materials[k, 14] > 0 for k in glass)
    a.materials[glass, 14] = 0
    rs2 = oracle_lib.RefScene(a)
    d2, i2, _ = rs2.render(1, rs2.params())
    assert np.array_equal(d, d2) and np.array_equal(i, i2)
    # ... while removing the diffuse wall's texture changes the image.
    wall = [k for k, m in enumerate(a.materials) if m[0] == 1 and m[14] > 0][0]
    a.materials[wall, 14] = 0
    rs3 = oracle_lib.RefScene(a)
    d3, _, _ = rs3.render(1, rs3.params())
    assert not np.array_equal(d, d3)


def test_fixture_script_is_committed_beside_the_fixtures():
    here = os.path.dirname(GOLDEN)
    assert os.path.exists(os.path.join(here, "make_image_fixtures.py"))
    # Expected arrays are plain uint8 npz entries (no pickled objects).
    with np.load(os.path.join(GOLDEN, "expected.npz"), allow_pickle=False) as z:
        assert len(z.files) == len(NAMES)
    assert struct.calcsize("<ii") == 8
