"""Pins the oracle's G-buffer and SVGF passes (oracle/rtref_svgf.inc) without
re-typing their formulas: closed forms, invariants and planted edits.

G-buffer (renderer.cuh:1446-1503, Shader/gbuffer.*): positions, ids,
barycentrics and normals against the scene's triangles; NDC of every hit
against the pixel centre (ties the LookAt/Perspective product to the pinhole
camera); motion against a float64 projection through the previous pinhole
camera; linear depth against the view-axis distance; fwidth against the quad
partners' depths; both faces seen.

SVGF (renderer.cuh:442-954): the first frame; a running mean on a static
G-buffer (alpha = 1/n); integer and half-pixel motion; disocclusion, depth and
normal rejection; the history cap; the variance pass's pass-through and
moment closed form; the a-trous pass as a plain B3 filter when the edge-stops
are disabled, its id edge-stop, its dilation step and its variance
propagation (sum w^2 v / (sum w)^2).
"""
import numpy as np
import pytest

from oracle import rtref

W, H = 64, 48


def uniform_g(w=W, h=H, mesh=1, mat=0, tri=0, depth=2.0, dz=0.01, motion=(0.0, 0.0), normal=(0.0, 0.0, 1.0)):
    g = rtref.GBuffer.empty(w, h)
    g.pos_tri[..., 3] = tri
    n = np.asarray(normal, np.float32)
    n = n / np.linalg.norm(n)
    g.nrm_mat[..., :3] = ((n * np.float32(0.5) + np.float32(0.5)) * np.float32(65535.0)).astype(np.uint32)
    g.nrm_mat[..., 3] = mat
    g.bary_mesh[..., 3] = mesh
    g.motion_depth[..., 0] = motion[0]
    g.motion_depth[..., 1] = motion[1]
    g.motion_depth[..., 2] = depth
    g.motion_depth[..., 3] = dz
    return g


def rgb(x):
    """(h, w) luminance-like scalar -> (h, w, 4) grey colour."""
    out = np.zeros(x.shape + (4,), np.float32)
    out[..., :3] = x[..., None]
    return out


def lum(c):
    c = c.astype(np.float64)
    return 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]


class History:
    """Feeds a temporal pass its own previous outputs."""

    def __init__(self, p, w=W, h=H):
        self.p = p
        self.hd = np.zeros((h, w, 4), np.float32)
        self.hi = np.zeros((h, w, 4), np.float32)
        self.md = np.zeros((h, w, 2), np.float32)
        self.mi = np.zeros((h, w, 2), np.float32)
        self.ln = np.zeros((h, w), np.int32)
        self.g = None

    def step(self, acc_d, acc_i, g, first=False, inv_spp=1.0):
        prev = self.g if self.g is not None else g
        out = rtref.svgf_temporal(first, self.p, inv_spp, acc_d, acc_i, self.hd, self.hi, self.md, self.mi,
                                  self.ln, g, prev)
        self.hd, self.hi, self.md, self.mi, self.ln = out
        self.g = g
        return out


# ------------------------------------------------------------------ temporal
def test_temporal_first_frame(oracle_lib):
    rng = np.random.default_rng(1)
    acc_d = rng.random((H, W, 4), dtype=np.float32) * 3
    acc_i = rng.random((H, W, 4), dtype=np.float32) * 3
    hd, hi, md, mi, ln = History(rtref.SvgfParams.default()).step(acc_d, acc_i, uniform_g(), first=True,
                                                                  inv_spp=np.float32(1) / np.float32(3))
    inv = np.float32(1) / np.float32(3)
    np.testing.assert_array_equal(hd[..., :3], acc_d[..., :3] * inv)
    np.testing.assert_array_equal(hi[..., :3], acc_i[..., :3] * inv)
    assert (hd[..., 3] == 1).all() and (hi[..., 3] == 1).all() and (ln == 1).all()
    ld = lum(acc_d * inv)
    np.testing.assert_allclose(md[..., 0], ld, rtol=1e-6)
    np.testing.assert_allclose(md[..., 1], ld * ld, rtol=1e-5)


def test_temporal_running_mean_on_static_gbuffer(oracle_lib):
    """Static camera, one surface: alpha = 1/n, so the history is the running
    mean of the frames and the moments are the mean of L and L^2."""
    rng = np.random.default_rng(2)
    p = rtref.SvgfParams.default()
    hist = History(p)
    g = uniform_g()
    xs = []
    for n in range(1, 9):
        x = rng.random((H, W), dtype=np.float32) * 2
        xs.append(x.astype(np.float64))
        hd, hi, md, mi, ln = hist.step(rgb(x), rgb(x * 0.5), g, first=(n == 1))
        inner = (slice(0, H - 1), slice(0, W - 1))  # the 2x2 footprint stays inside the image
        mean = np.mean(xs, axis=0)
        np.testing.assert_allclose(hd[..., 0][inner], mean[inner], rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(hi[..., 1][inner], 0.5 * mean[inner], rtol=2e-6, atol=1e-7)
        # GetLastFramePos accepts 0.5 <= pos < size - 0.5 (renderer.cuh:450): the last
        # column and row never reproject, every other pixel extends its history.
        assert (ln[inner] == n).all(), n
        assert (ln[-1, :] == 1).all() and (ln[:, -1] == 1).all()
        l_mean = np.mean([lum(rgb(x.astype(np.float32))) for x in xs], axis=0)
        l2_mean = np.mean([lum(rgb(x.astype(np.float32))) ** 2 for x in xs], axis=0)
        np.testing.assert_allclose(md[..., 0][inner], l_mean[inner], rtol=1e-5)
        np.testing.assert_allclose(md[..., 1][inner], l2_mean[inner], rtol=1e-5)
        if n > 1:
            np.testing.assert_allclose(hd[..., 3][inner], (l2_mean - l_mean ** 2)[inner], rtol=1e-3, atol=1e-5)


def test_temporal_history_cap(oracle_lib):
    p = rtref.SvgfParams.default(max_history=3)
    hist = History(p)
    g = uniform_g()
    rng = np.random.default_rng(3)
    ref = None
    for n in range(1, 8):
        x = rng.random((H, W), dtype=np.float32)
        hd, _, _, _, ln = hist.step(rgb(x), rgb(x), g, first=(n == 1))
        a = 1.0 / min(n, 3)
        ref = x.astype(np.float64) if ref is None else a * x + (1 - a) * ref
        assert (ln[:-1, :-1] == min(n, 3)).all()
        np.testing.assert_allclose(hd[:-1, :-1, 0], ref[:-1, :-1], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("shift", [2.0, 0.5, -1.0])
def test_temporal_reprojection_follows_motion(oracle_lib, shift):
    """motion.x (NDC) = 2 * shift / W moves the history tap `shift` pixels to the
    left; half a pixel blends the two neighbours 50/50."""
    p = rtref.SvgfParams.default()
    hist = History(p)
    ramp = np.tile(np.arange(W, dtype=np.float32), (H, 1))
    hist.step(rgb(ramp), rgb(ramp), uniform_g(), first=True)
    cur = rng_frame = np.full((H, W), 7.0, np.float32)
    g2 = uniform_g(motion=(2.0 * shift / W, 0.0))
    hd, _, _, _, ln = hist.step(rgb(cur), rgb(cur), g2)
    src = np.arange(W) - shift
    ok = (src >= 0) & (src <= W - 2) & (np.arange(W) <= W - 2)
    expect = 0.5 * 7.0 + 0.5 * src
    rows = slice(0, H - 1)
    np.testing.assert_allclose(hd[rows][:, ok, 0], np.broadcast_to(expect[ok], (H - 1, ok.sum())), rtol=1e-6)
    assert (ln[rows][:, ok] == 2).all()
    del rng_frame


def test_temporal_disocclusion_resets(oracle_lib):
    p = rtref.SvgfParams.default()
    hist = History(p)
    g_prev = uniform_g()
    g_prev.bary_mesh[:, :32, 3] = 2  # another mesh covered the left half last frame
    hist.step(rgb(np.ones((H, W), np.float32)), rgb(np.ones((H, W), np.float32)), g_prev, first=True)
    cur = np.full((H, W), 3.0, np.float32)
    hd, hi, md, mi, ln = hist.step(rgb(cur), rgb(cur), uniform_g())
    reset = np.s_[:, :31]  # every 3x3 fallback tap is on the other mesh
    assert (ln[reset] == 1).all()
    assert (hd[reset][..., 0] == 3.0).all() and (hd[reset][..., 3] == 1.0).all()
    kept = np.s_[:H - 1, 33:W - 1]
    assert (ln[kept] == 2).all()
    np.testing.assert_allclose(hd[kept][..., 0], 2.0, rtol=1e-6)


@pytest.mark.parametrize("depth,valid", [(2.2, True), (3.0, False)])
def test_temporal_depth_threshold(oracle_lib, depth, valid):
    """|d - d'| <= max(dz, dz') (px + 1) + 0.1 max(d, d') (renderer.cuh:491)."""
    hist = History(rtref.SvgfParams.default())
    hist.step(rgb(np.ones((H, W), np.float32)), rgb(np.ones((H, W), np.float32)), uniform_g(depth=2.0), first=True)
    _, _, _, _, ln = hist.step(rgb(np.ones((H, W), np.float32)), rgb(np.ones((H, W), np.float32)),
                               uniform_g(depth=depth))
    assert (ln[:-1, :-1] == (2 if valid else 1)).all()


@pytest.mark.parametrize("angle,valid", [(20.0, True), (30.0, False)])
def test_temporal_normal_threshold(oracle_lib, angle, valid):
    hist = History(rtref.SvgfParams.default())
    hist.step(rgb(np.ones((H, W), np.float32)), rgb(np.ones((H, W), np.float32)), uniform_g(), first=True)
    a = np.deg2rad(angle)
    _, _, _, _, ln = hist.step(rgb(np.ones((H, W), np.float32)), rgb(np.ones((H, W), np.float32)),
                               uniform_g(normal=(np.sin(a), 0.0, np.cos(a))))
    assert (ln[:-1, :-1] == (2 if valid else 1)).all()


# ------------------------------------------------------------------ variance
def test_variance_long_history_passes_through(oracle_lib):
    rng = np.random.default_rng(4)
    d = rng.random((H, W, 4), dtype=np.float32)
    i = rng.random((H, W, 4), dtype=np.float32)
    m = rng.random((H, W, 2), dtype=np.float32)
    ln = np.full((H, W), 4, np.int32)
    od, oi = rtref.svgf_variance(rtref.SvgfParams.default(), d, i, m, m, ln, uniform_g())
    np.testing.assert_array_equal(od, d)
    np.testing.assert_array_equal(oi, i)


@pytest.mark.parametrize("hl", [1, 2, 3])
def test_variance_spatial_estimate_closed_form(oracle_lib, hl):
    """Constant colour and moments (m1, m2): colour kept, variance = (m2 - m1^2) * 4 / hl."""
    c = np.full((H, W), 0.8, np.float32)
    m = np.zeros((H, W, 2), np.float32)
    m[..., 0], m[..., 1] = 0.5, 0.4
    ln = np.full((H, W), hl, np.int32)
    od, oi = rtref.svgf_variance(rtref.SvgfParams.default(), rgb(c), rgb(c), m, m, ln, uniform_g())
    np.testing.assert_allclose(od[..., :3], 0.8, rtol=1e-6)
    np.testing.assert_allclose(od[..., 3], (0.4 - 0.25) * 4 / hl, rtol=1e-5)
    np.testing.assert_allclose(oi[..., 3], (0.4 - 0.25) * 4 / hl, rtol=1e-5)


def test_variance_isolated_surface_keeps_own_moments(oracle_lib):
    rng = np.random.default_rng(5)
    c = rng.random((H, W), dtype=np.float32) + 0.1
    m = rng.random((H, W, 2), dtype=np.float32)
    m[..., 1] += m[..., 0] ** 2
    g = uniform_g()
    g.bary_mesh[..., 3] = np.arange(W * H, dtype=np.uint16).reshape(H, W) + 1  # every pixel its own mesh
    ln = np.ones((H, W), np.int32)
    od, _ = rtref.svgf_variance(rtref.SvgfParams.default(), rgb(c), rgb(c), m, m, ln, g)
    np.testing.assert_allclose(od[..., 0], c, rtol=1e-6)
    np.testing.assert_allclose(od[..., 3], np.maximum(0, m[..., 1] - m[..., 0] ** 2) * 4, rtol=1e-4, atol=1e-6)


# ------------------------------------------------------------------ a-trous
B3 = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])


def b3_filter(x, step, v=None):
    """Edge-normalised 5x5 B3 filter at dilation `step`, float64; with v the
    propagated variance sum(w^2 v) / (sum w)^2."""
    h, w = x.shape
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    vnum = np.zeros((h, w))
    for b in range(-2, 3):
        for a in range(-2, 3):
            k = B3[a + 2] * B3[b + 2]
            dy, dx = b * step, a * step
            ys, xs = np.arange(h) + dy, np.arange(w) + dx
            my, mx = (ys >= 0) & (ys < h), (xs >= 0) & (xs < w)
            m = np.outer(my, mx)
            src = np.zeros((h, w))
            src[np.ix_(my, mx)] = x[np.ix_(ys[my], xs[mx])]
            num += m * k * src
            den += m * k
            if v is not None:
                vs = np.zeros((h, w))
                vs[np.ix_(my, mx)] = v[np.ix_(ys[my], xs[mx])]
                vnum += m * k * k * vs
    return num / den, (vnum / den ** 2 if v is not None else None)


@pytest.mark.parametrize("step", [1, 2, 4, 8])
def test_atrous_is_b3_filter_without_edge_stops(oracle_lib, step):
    """sigma_light huge, one surface, long history: the edge-stopping weights are
    all equal and the pass is the plain B3 a-trous filter."""
    rng = np.random.default_rng(6 + step)
    x = rng.random((H, W), dtype=np.float32)
    v = rng.random((H, W), dtype=np.float32) * 0.1
    c = rgb(x)
    c[..., 3] = v
    p = rtref.SvgfParams.default(sigma_light=1e20)
    ln = np.full((H, W), 4, np.int32)
    od, oi = rtref.svgf_atrous(step, p, c, c, ln, uniform_g())
    ex, ev = b3_filter(x.astype(np.float64), step, v.astype(np.float64))
    np.testing.assert_allclose(od[..., 0], ex, rtol=2e-5)
    np.testing.assert_allclose(oi[..., 2], ex, rtol=2e-5)
    np.testing.assert_allclose(od[..., 3], ev, rtol=2e-4)


def test_atrous_constant_image_is_fixed_point(oracle_lib):
    c = rgb(np.full((H, W), 0.37, np.float32))
    c[..., 3] = 0.0
    ln = np.full((H, W), 2, np.int32)  # bootstrap branch
    for step in (1, 2, 4, 8, 16):
        od, oi = rtref.svgf_atrous(step, rtref.SvgfParams.default(), c, c, ln, uniform_g())
        np.testing.assert_allclose(od[..., :3], 0.37, rtol=1e-6)
        assert (od[..., 3] == 0).all()


def test_atrous_stops_at_mesh_and_material_ids(oracle_lib):
    x = np.where(np.arange(W) < 30, 1.0, 5.0).astype(np.float32)
    c = rgb(np.tile(x, (H, 1)))
    c[..., 3] = 0.5
    g = uniform_g()
    g.bary_mesh[:, 30:, 3] = 2
    g.nrm_mat[:20, :, 3] = 3  # a material split across the other axis
    p = rtref.SvgfParams.default(sigma_light=1e20)
    ln = np.full((H, W), 4, np.int32)
    od, _ = rtref.svgf_atrous(1, p, c, c, ln, g)
    np.testing.assert_allclose(od[..., 0], c[..., 0], rtol=1e-6)


def test_atrous_dilation_reaches_only_multiples_of_step(oracle_lib):
    base = np.full((H, W), 0.5, np.float32)
    spike = base.copy()
    spike[24, 32] = 50.0
    p = rtref.SvgfParams.default(sigma_light=1e20)
    ln = np.full((H, W), 4, np.int32)
    g = uniform_g()
    step = 4
    a, _ = rtref.svgf_atrous(step, p, rgb(base), rgb(base), ln, g)
    b, _ = rtref.svgf_atrous(step, p, rgb(spike), rgb(spike), ln, g)
    changed = np.argwhere(a[..., 0] != b[..., 0])
    for y, x in changed:
        if (y, x) == (24, 32):
            continue
        assert (y - 24) % step == 0 and (x - 32) % step == 0 and abs(y - 24) <= 2 * step and abs(x - 32) <= 2 * step
    assert len(changed) == 25


def test_atrous_isolated_pixel_keeps_colour_and_variance(oracle_lib):
    rng = np.random.default_rng(9)
    c = rgb(rng.random((H, W), dtype=np.float32) + 0.05)
    c[..., 3] = rng.random((H, W), dtype=np.float32)
    g = uniform_g()
    g.bary_mesh[..., 3] = np.arange(W * H, dtype=np.uint16).reshape(H, W) + 1
    ln = np.full((H, W), 4, np.int32)
    od, _ = rtref.svgf_atrous(2, rtref.SvgfParams.default(), c, c, ln, g)
    np.testing.assert_allclose(od[..., 0], c[..., 0], rtol=1e-6)
    np.testing.assert_allclose(od[..., 3], c[..., 3], rtol=1e-5)


# ------------------------------------------------------------------ G-buffer
@pytest.fixture(scope="module")
def rt_scene(scene_dir, oracle_lib):
    a = rtref.parse_scene(scene_dir["cornell_spheres_rt"])
    return a, rtref.RefScene(a)


def _cam(s, orbit):
    keep = s.orbit()
    s.set_orbit(orbit)
    c = s.camera().astype(np.float64)
    s.set_orbit(keep)
    return c  # position, left_down, du, dv


def _project(cam, p, w, h):
    """Pixel coordinates of world points through a pinhole camera (float64):
    solve ld + x du + y dv = pos + lam (p - pos)."""
    pos, ld, du, dv = cam
    A = np.stack([np.broadcast_to(du, p.shape), np.broadcast_to(dv, p.shape), -(p - pos)], axis=-1)
    sol = np.linalg.solve(A, np.broadcast_to(pos - ld, p.shape)[..., None])[..., 0]
    return sol[..., 0], sol[..., 1]


def test_gbuffer_surface_matches_scene_triangles(rt_scene):
    a, s = rt_scene
    g = s.gbuffer()
    hit = g.bary_mesh[..., 3] > 0
    assert hit.mean() > 0.8
    mesh = g.bary_mesh[..., 3][hit].astype(np.int64) - 1
    starts = np.array([np.flatnonzero(a.mesh == m)[0] for m in range(len(a.mesh_material))])
    tri = starts[mesh] + g.pos_tri[..., 3][hit].astype(np.int64)
    assert (a.mesh[tri] == mesh).all()
    assert (g.nrm_mat[..., 3][hit] == a.mesh_material[mesh]).all()
    bary = g.bary_mesh[..., :3][hit].astype(np.float64) / 65535.0
    assert np.abs(bary.sum(-1) - 1).max() < 1e-4
    v = a.pos[tri].reshape(-1, 3, 3).astype(np.float64)
    n = a.nrm[tri].reshape(-1, 3, 3).astype(np.float64)
    p = np.einsum("nk,nkc->nc", bary, v)
    np.testing.assert_allclose(g.pos_tri[..., :3][hit], p, atol=2e-4)
    ni = np.einsum("nk,nkc->nc", bary, n)
    ni /= np.linalg.norm(ni, axis=-1, keepdims=True)
    dec = g.nrm_mat[..., :3][hit].astype(np.float64) / 65535.0 * 2 - 1
    np.testing.assert_allclose(dec, ni, atol=2e-4)
    assert (g.nrm_mat[~hit] == 0).all() and (g.motion_depth[~hit] == 0).all()


def test_gbuffer_projection_matches_pinhole_camera(rt_scene):
    """VP maps each pixel-centre hit to that pixel's NDC; linear depth is the
    view-axis distance; motion is NDC(current) - NDC(previous pinhole)."""
    a, s = rt_scene
    o = s.orbit()
    prev = rtref.Orbit(o.target, o.distance, o.theta_deg - 3.0, o.phi_deg + 4.0)
    g = s.gbuffer(prev)
    hit = g.bary_mesh[..., 3] > 0
    jj, ii = np.nonzero(hit)
    p = g.pos_tri[..., :3][hit].astype(np.float64)
    cam = s.camera().astype(np.float64)
    pos = cam[0]
    lo, hi = a.pos.reshape(-1, 3).min(0), a.pos.reshape(-1, 3).max(0)
    far = np.linalg.norm(hi - lo) + np.linalg.norm(pos - (lo + hi) / 2)
    vp = s.view_projection(o, 0.1, far).astype(np.float64)
    clip = np.concatenate([p, np.ones((len(p), 1))], 1) @ vp.T
    np.testing.assert_allclose(clip[:, 0] / clip[:, 3], 2 * (ii + 0.5) / a.width - 1, atol=2e-5)
    np.testing.assert_allclose(clip[:, 1] / clip[:, 3], 2 * (jj + 0.5) / a.height - 1, atol=2e-5)
    fwd = np.asarray(o.target, np.float64) - pos
    fwd /= np.linalg.norm(fwd)
    np.testing.assert_allclose(g.motion_depth[..., 2][hit], (p - pos) @ fwd, rtol=1e-4)
    x, y = _project(_cam(s, prev), p, a.width, a.height)
    mx = 2 * (ii + 0.5) / a.width - 2 * x / a.width
    my = 2 * (jj + 0.5) / a.height - 2 * y / a.height
    np.testing.assert_allclose(g.motion_depth[..., 0][hit], mx, atol=5e-5)
    np.testing.assert_allclose(g.motion_depth[..., 1][hit], my, atol=5e-5)
    g0 = s.gbuffer()
    assert (g0.motion_depth[..., :2] == 0).all()


def test_gbuffer_fwidth_matches_quad_partners(rt_scene):
    """Inside a triangle fwidth(depth) = |d(i^1, j) - d| + |d(i, j^1) - d|."""
    a, s = rt_scene
    g = s.gbuffer()
    d = g.motion_depth[..., 2].astype(np.float64)
    key = g.bary_mesh[..., 3].astype(np.int64) * 1_000_000 + g.pos_tri[..., 3].astype(np.int64)
    jj, ii = np.mgrid[0:a.height, 0:a.width]
    px, py = ii ^ 1, jj ^ 1
    ok = (px < a.width) & (py < a.height) & (g.bary_mesh[..., 3] > 0)
    pxc, pyc = np.minimum(px, a.width - 1), np.minimum(py, a.height - 1)
    ok &= (key[jj, pxc] == key) & (key[pyc, ii] == key)
    assert ok.sum() > 0.3 * a.width * a.height
    expect = np.abs(d[jj, pxc] - d) + np.abs(d[pyc, ii] - d)
    np.testing.assert_allclose(g.motion_depth[..., 3][ok], expect[ok], rtol=2e-3, atol=1e-4)


def test_gbuffer_sees_back_faces(rt_scene):
    """From outside the box the rasterised G-buffer sees the wall's back face;
    the path tracer's culling traversal looks through it."""
    a, s = rt_scene
    o = s.orbit()
    keep = rtref.Orbit(o.target, o.distance, o.theta_deg, o.phi_deg)
    outside = rtref.Orbit(o.target, o.distance * 1.2, o.theta_deg, o.phi_deg + 180.0)
    s.set_orbit(outside)
    try:
        g = s.gbuffer()
        pos, ld, du, dv = s.camera()
        j, i = a.height // 2, a.width // 2
        pp = ld + (i + 0.5) * du + (j + 0.5) * dv
        dirn = (pp - pos) / np.linalg.norm(pp - pos)
        ray = np.concatenate([pos, [0.0], dirn, [1e8]]).astype(np.float32)[None]
        t, _, _, prim, _ = s.trace(ray)
        assert g.bary_mesh[j, i, 3] > 0
        hit_t = np.linalg.norm(g.pos_tri[j, i, :3] - pos)
        assert prim[0] < 0 or hit_t < t[0] - 1e-3
    finally:
        s.set_orbit(keep)
