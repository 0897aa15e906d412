"""Image files (rt_save_image): the PNG holds exactly the gathered colour of
the accumulation or of the last denoised frame (GatherKernel, renderer.cuh:
956-993, as UI::SaveScreenshot writes it, Src/Utils/ui.cpp:370-380), and the
PFM the float colour, in PFM's bottom-first row order."""
import numpy as np
import pytest

import cudaraytracer_amd as rt
from oracle.rtref_image import read_texture

pytestmark = pytest.mark.gpu


def _read_pfm(path):
    data = open(path, "rb").read()
    head, rest = data.split(b"\n", 1)
    dims, rest = rest.split(b"\n", 1)
    scale, rest = rest.split(b"\n", 1)
    assert head == b"PF" and float(scale) < 0  # colour, little-endian
    w, h = map(int, dims.split())
    return np.frombuffer(rest, "<f4", w * h * 3).reshape(h, w, 3)


def test_save_png_and_pfm(scene_dir, tmp_path):
    r = rt.Renderer(rt.Scene(scene_dir["cornell_spheres_rt"]))
    with pytest.raises(rt.RtError):
        r.save_image(rt.RT_IMAGE_PNG_DENOISED, str(tmp_path / "none.png"))
    r.render(4)
    r.save_image(rt.RT_IMAGE_PNG_COLOR, str(tmp_path / "color.png"))
    r.save_image(rt.RT_IMAGE_PFM_MEAN, str(tmp_path / "mean.pfm"))
    png = read_texture(str(tmp_path / "color.png"))
    assert np.array_equal(png, r.read_rgba8())
    mean = r.read_accum(rt.RT_ACCUM_MEAN)
    assert np.array_equal(_read_pfm(str(tmp_path / "mean.pfm")), mean.reshape(r.height, r.width, 4)[..., :3])
    # The denoised frame.
    r.render_frame(1)
    r.save_image(rt.RT_IMAGE_PNG_DENOISED, str(tmp_path / "den.png"))
    r.save_image(rt.RT_IMAGE_PFM_DENOISED, str(tmp_path / "den.pfm"))
    den = r.read_frame(rt.RT_FRAME_DENOISED).reshape(r.height, r.width, 4)
    assert np.array_equal(_read_pfm(str(tmp_path / "den.pfm")), den[..., :3])
    want = np.clip(np.power(den[::-1, :, :3].astype(np.float32), np.float32(1 / 2.2)) * np.float32(255), 0, 255)
    got = read_texture(str(tmp_path / "den.png"))
    assert np.abs(got[..., :3].astype(int) - want.astype(np.uint8).astype(int)).max() <= 1  # host pow vs numpy pow
    assert (got[..., 3] == 255).all()
    with pytest.raises(rt.RtError):
        r.save_image(7, str(tmp_path / "bad.png"))


def test_gather_modes(scene_dir):
    """GatherKernel's render-target views (renderer.cuh:956-1029) of the last
    frame's G-buffer, restated from the G-buffer planes (themselves bit-exact
    against the oracle in test_gpu_svgf.py), top row first."""
    scene = rt.Scene(scene_dir["cornell_spheres_rt"])
    info = scene.info()
    r = rt.Renderer(scene)
    with pytest.raises(rt.RtError):
        r.read_gather(rt.RT_GATHER_DEPTH)
    r.render(2)
    assert np.array_equal(r.read_gather(rt.RT_GATHER_COLOR), r.read_rgba8())
    r.render_frame(1)
    r.render_frame(1)
    f32 = np.float32
    flip = lambda a: a.reshape(r.height, r.width, -1)[::-1]
    # Normal: u16 / 65535 * 255, truncated.
    nm = flip(r.read_frame(rt.RT_FRAME_G_NORMAL_MAT)).astype(np.float32)
    want = np.clip((nm[..., :3] / f32(65535)) * f32(255), 0, 255).astype(np.uint8)
    assert np.array_equal(r.read_gather(rt.RT_GATHER_NORMAL)[..., :3], want)
    # Mesh id colours.
    ids = flip(r.read_frame(rt.RT_FRAME_G_BARY_MESH))[..., 3].astype(np.int64)
    want = np.stack([(ids * 37) % 256, (ids * 57) % 256, (ids * 97) % 256], -1).astype(np.uint8)
    got = r.read_gather(rt.RT_GATHER_ID)
    assert np.array_equal(got[..., :3], want) and (got[..., 3] == 255).all()
    assert (ids == 0).any() and len(np.unique(ids)) > 3  # background and several meshes in view
    # Position: scaled by the scene bounds.
    lo, hi = np.array(info.bounds_lo, f32), np.array(info.bounds_hi, f32)
    pos = flip(r.read_frame(rt.RT_FRAME_G_POSITION_TRI))[..., :3]
    want = np.clip((pos - lo) / (hi - lo) * f32(255), 0, 255).astype(np.uint8)
    assert np.array_equal(r.read_gather(rt.RT_GATHER_POSITION)[..., :3], want)
    # Depth: (z - 0.1) / (far - 0.1), far = scene diagonal + camera distance to the centre.
    length = lambda v: np.sqrt(v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]))
    cam = np.array(info.camera_position, f32)
    far = length(hi - lo) + length(cam - (lo + hi) * f32(0.5))
    z = flip(r.read_frame(rt.RT_FRAME_G_MOTION_DEPTH))[..., 2]
    want = np.clip((z - f32(0.1)) / (far - f32(0.1)) * f32(255), 0, 255).astype(np.uint8)
    got = r.read_gather(rt.RT_GATHER_DEPTH)
    assert np.abs(got[..., 0].astype(int) - want.astype(int)).max() <= 1
    assert (got[..., 0] == got[..., 1]).all() and (got[..., 0] == got[..., 2]).all()
    # Denoised colour: gamma gather of the denoised frame.
    den = flip(r.read_frame(rt.RT_FRAME_DENOISED))[..., :3]
    want = np.clip(np.power(den, f32(1 / 2.2)) * f32(255), 0, 255).astype(np.uint8)
    assert np.abs(r.read_gather(rt.RT_GATHER_DENOISED)[..., :3].astype(int) - want.astype(int)).max() <= 1
    with pytest.raises(rt.RtError):
        r.read_gather(9)


def test_read_denoised_and_gbuffer_entry_points(scene_dir):
    import ctypes

    r = rt.Renderer(rt.Scene(scene_dir["cornell_spheres_rt"]))
    r.render_frame(1)
    lib = rt.lib()
    den = np.zeros((r.height, r.width, 4), np.float32)
    assert lib.rt_read_denoised(r._h, den.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    assert np.array_equal(den.reshape(-1), r.read_frame(rt.RT_FRAME_DENOISED).reshape(-1))
    planes = [rt.RT_FRAME_G_POSITION_TRI, rt.RT_FRAME_G_NORMAL_MAT, rt.RT_FRAME_G_BARY_MESH,
              rt.RT_FRAME_G_MOTION_DEPTH]
    for c, which in enumerate(planes):
        want = r.read_frame(which)
        got = np.zeros_like(want)
        assert lib.rt_read_gbuffer(r._h, c, got.ctypes.data_as(ctypes.c_void_p)) == 0
        assert np.array_equal(got, want)
    assert lib.rt_read_gbuffer(r._h, 4, den.ctypes.data_as(ctypes.c_void_p)) != 0
