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
