"""GPU parity: the HIP kernels against the CPU oracle on the same inputs.

Integer / index results (hit primitive, traversal counts) and the FP32
radiance sums are required to be bit-identical: both sides run the same FP32
operation sequence with contraction off and the same counter-based RNG."""
import numpy as np
import pytest

import cudaraytracer_amd as rt

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


@pytest.mark.parametrize("op,lo,hi", [(2, -20.0, 20.0), (3, -20.0, 20.0), (4, -1.0, 1.0), (5, -90.0, 90.0),
                                      (6, 1e-30, 1e30), (7, -5.0, 5.0), (10, 0.0, 1.0)])
def test_device_transcendentals_bit_exact(oracle_lib, op, lo, hi):
    rng = np.random.default_rng(op)
    if op == 6:
        a = np.exp(rng.uniform(np.log(lo), np.log(hi), 20000)).astype(np.float32)
    else:
        a = rng.uniform(lo, hi, 20000).astype(np.float32)
    b = rng.uniform(0, 300, 20000).astype(np.float32) if op == 10 else rng.uniform(-5, 5, 20000).astype(np.float32)
    if op == 10:
        a[:100] = 0.0  # 0^y
    got = rt.device_math(op, a, b)
    want = np.array([oracle_lib.math(op, float(x), float(y)) for x, y in zip(a, b)], np.float32)
    assert np.array_equal(_bits(got), _bits(want))


def test_device_ieee_ops_bit_exact():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal(1 << 18) * 1e3).astype(np.float32)
    b = (rng.standard_normal(1 << 18) + 0.01).astype(np.float32)
    assert np.array_equal(_bits(rt.device_math(0, a, b)), _bits(a / b))
    assert np.array_equal(_bits(rt.device_math(1, np.abs(a), b)), _bits(np.sqrt(np.abs(a))))
    assert np.array_equal(_bits(rt.device_math(9, b, b)), _bits(np.float32(1) / b))


def test_device_philox_matches_oracle(oracle_lib):
    rng = np.random.default_rng(5)
    n = 4096
    a = rng.integers(0, 2**31, n).astype(np.int32).view(np.float32)
    b = rng.integers(0, 2**31, n).astype(np.int32).view(np.float32)
    got = rt.device_math(8, a, b).view(np.uint32)
    for i in range(0, n, 97):
        c = int(a[i:i + 1].view(np.uint32)[0])
        k = int(b[i:i + 1].view(np.uint32)[0])
        want = oracle_lib.philox([c, c ^ 0x9E3779B9, i, 7], [k, 0x5EEDC0DE])
        assert got[i] == want[i & 3]


def _camera_and_random_rays(scene_arrays, rs, n_random=20000, seed=3):
    cam = rs.camera()
    pos, ld, du, dv = cam
    w, h = scene_arrays.width, scene_arrays.height
    rng = np.random.default_rng(seed)
    i = rng.integers(0, w, n_random)
    j = rng.integers(0, h, n_random)
    p = ld[None] + (i[:, None] + rng.random((n_random, 1))) * du[None] + (j[:, None] + rng.random((n_random, 1))) * dv[None]
    d = p - pos[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cam_rays = np.concatenate([np.repeat(pos[None], n_random, 0), np.zeros((n_random, 1)), d,
                               np.full((n_random, 1), 1e8)], axis=1)
    lo, hi = scene_arrays.pos.reshape(-1, 3).min(0), scene_arrays.pos.reshape(-1, 3).max(0)
    o = rng.uniform(lo, hi, (n_random, 3))
    d2 = rng.standard_normal((n_random, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    d2[:16] = np.eye(3)[np.arange(16) % 3] * np.where(np.arange(16) % 2, 1, -1)[:, None]  # axis-parallel
    tmin = rng.choice([0.0, 1e-6, 1e-4], (n_random, 1))
    tmax = rng.choice([1e8, 0.5, 2.0], (n_random, 1))
    inner = np.concatenate([o, tmin, d2, tmax], axis=1)
    return np.concatenate([cam_rays, inner]).astype(np.float32)


@pytest.mark.parametrize("name", ["cornell_small", "cornell_spheres", "terrain_small"])
def test_traversal_bit_exact(scene_dir, oracle_lib, name):
    s = rt.Scene(scene_dir[name])
    a = oracle_lib.parse_scene(scene_dir[name])
    rs = oracle_lib.RefScene(a)
    rays = _camera_and_random_rays(a, rs)
    gt, gb1, gb2, gp, gc = s.trace(rays)
    rt_, rb1, rb2, rp, rc = rs.trace(rays)
    assert np.array_equal(gp, rp)
    hit = rp >= 0
    assert hit.mean() > 0.15
    assert np.array_equal(_bits(gt), _bits(rt_))
    assert np.array_equal(_bits(gb1[hit]), _bits(rb1[hit])) and np.array_equal(_bits(gb2[hit]), _bits(rb2[hit]))
    assert np.array_equal(gc, rc)  # same nodes tested and triangles tested per ray


def _gpu_render(path, spp, **kw):
    s = rt.Scene(path)
    r = rt.Renderer(s, **kw)
    r.render(spp)
    d = r.read_accum(rt.RT_ACCUM_DIRECT).reshape(-1, 4)
    i = r.read_accum(rt.RT_ACCUM_INDIRECT).reshape(-1, 4)
    st = r.stats()
    return d, i, st, r


@pytest.mark.parametrize("name,spp", [("cornell_small", 4), ("cornell_spheres_128", 4), ("terrain_small", 4),
                                      ("terrain_small_burley", 4), ("cornell_sss_walk", 4)])
def test_render_bit_exact_full_image(scene_dir, oracle_lib, name, spp):
    d, i, st, _ = _gpu_render(scene_dir[name], spp)
    a = oracle_lib.parse_scene(scene_dir[name])
    rs = oracle_lib.RefScene(a)
    rd, ri, rst = rs.render(spp, rs.params())
    assert np.array_equal(_bits(d[:, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[:, :3]), _bits(ri[:, :3]))
    assert st.extension_rays == rst[0]
    assert st.shadow_rays == rst[1]
    assert st.probe_rays == rst[2]  # Burley probe rays / random-walk segments
    if name == "cornell_sss_walk":
        assert rst[2] > 10 * 4 * 96 * 96 // 100  # the walks are exercised


def test_render_reference_semantics_variants(scene_dir, oracle_lib):
    """Roulette from bounce 0 (the reference's rule) and flushed survivors."""
    path = scene_dir["cornell_spheres_128"]
    a = oracle_lib.parse_scene(path)
    rs = oracle_lib.RefScene(a)
    for kw in ({"rr_start_depth": 0}, {"flush_survivors": True, "max_bounces": 3}):
        d, i, _, _ = _gpu_render(path, 2, **kw)
        rd, ri, _ = rs.render(2, rs.params(**kw))
        assert np.array_equal(_bits(d[:, :3]), _bits(rd[:, :3]))
        assert np.array_equal(_bits(i[:, :3]), _bits(ri[:, :3]))


def test_env_alias_sampling_bit_exact(scene_dir, oracle_lib):
    """The opt-in alias-table environment sampler against the oracle's, and
    against the CDF sampler in expectation (same texel distribution and pdf)."""
    path = scene_dir["terrain_small"]
    a = oracle_lib.parse_scene(path)
    rs = oracle_lib.RefScene(a)
    d, i, st, _ = _gpu_render(path, 4, env_alias=True)
    rd, ri, rst = rs.render(4, rs.params(env_alias=True))
    assert np.array_equal(_bits(d[:, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[:, :3]), _bits(ri[:, :3]))
    assert st.shadow_rays == rst[1]
    # Same estimator, different samples: the means over the image agree.
    dc, ic, _, _ = _gpu_render(path, 64)
    da, ia, _, _ = _gpu_render(path, 64, env_alias=True)
    mc, ma = (dc + ic)[:, :3].mean(), (da + ia)[:, :3].mean()
    assert abs(mc - ma) < 0.05 * mc


def test_graph_replays_match_eager_launches(scene_dir):
    """Every pass after the first of a shape is a replay of the captured graph;
    replays interleaved with other eager work and synchronisations must give
    the eager launches' results, pass after pass."""
    path = scene_dir["cornell_spheres_128"]
    s = rt.Scene(path)
    eager = rt.Renderer(s, batch_spp=1, graphs=False)
    graph = rt.Renderer(s, batch_spp=1)
    other = rt.Renderer(rt.Scene(scene_dir["cornell_small"]), batch_spp=2)
    for k in range(5):
        eager.render(1)
        graph.render(1)
        other.render(2)  # unrelated eager work between the replays
        other.synchronize()
        for which in (rt.RT_ACCUM_DIRECT, rt.RT_ACCUM_INDIRECT):
            assert np.array_equal(_bits(graph.read_accum(which)), _bits(eager.read_accum(which))), (k, which)
    assert graph.stats().extension_rays == eager.stats().extension_rays


def test_batching_does_not_change_results(scene_dir):
    path = scene_dir["cornell_spheres_128"]
    d1, i1, _, _ = _gpu_render(path, 6, batch_spp=1)
    d6, i6, _, _ = _gpu_render(path, 6, batch_spp=6)
    d4, i4, _, _ = _gpu_render(path, 6, batch_spp=4)
    assert np.array_equal(d1, d6) and np.array_equal(i1, i6)
    assert np.array_equal(d4, d6) and np.array_equal(i4, i6)
    s = rt.Scene(path)
    r = rt.Renderer(s)
    r.render(2)
    r.render(4)
    assert np.array_equal(r.read_accum(rt.RT_ACCUM_DIRECT).reshape(-1, 4), d6)


def test_full_size_workload_matches_oracle_on_sampled_pixels(scene_dir, oracle_lib):
    """512x512, 64 spp, depth 12: every sampled pixel's sums are bit-identical."""
    path = scene_dir["cornell_spheres"]
    d, i, st, r = _gpu_render(path, 64)
    img = r.read_accum(rt.RT_ACCUM_MEAN)
    assert np.isfinite(img).all()
    a = oracle_lib.parse_scene(path)
    rs = oracle_lib.RefScene(a)
    pix = np.random.default_rng(9).choice(a.width * a.height, 600, replace=False).astype(np.int32)
    rd, ri, _ = rs.render(64, rs.params(), pixels=pix)
    assert np.array_equal(_bits(d[pix, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[pix, :3]), _bits(ri[:, :3]))
    assert st.spp_done == 64


def test_rgba8_readback(scene_dir):
    _, _, _, r = _gpu_render(scene_dir["cornell_small"], 4)
    mean = r.read_accum(rt.RT_ACCUM_MEAN)
    img8 = r.read_rgba8()
    want = np.clip(np.power(mean[..., :3], np.float32(1 / 2.2)) * 255.0, 0, 255).astype(np.uint8)[::-1]
    assert np.abs(img8[..., :3].astype(int) - want.astype(int)).max() <= 1
    assert (img8[..., 3] == 255).all()


def test_traversal_counters(scene_dir):
    s = rt.Scene(scene_dir["cornell_spheres_128"])
    r = rt.Renderer(s)
    r.render(2, count_traversal=True)
    st = r.stats()
    assert st.extension_node_visits > st.extension_rays
    assert st.extension_tri_tests > 0 and st.shadow_node_visits > 0
    r.reset()
    r.reset_stats()
    r.render(2)
    st2 = r.stats()
    assert st2.extension_node_visits == 0 and st2.extension_rays == st.extension_rays


@pytest.mark.parametrize("name", ["cornell_spheres_128", "textured_cornell", "cornell_sss_walk"])
def test_converged_image_within_rmse_of_fp64_oracle(scene_dir, oracle_lib, name):
    """The GPU image against the oracle's FP64 twin (the same estimator and
    random numbers with FP64 shading): per pixel, 97 % within rtol 1e-3 /
    atol 1e-4 (rounding differences can flip a rare branch), and an image
    RMSE below 5 % of the mean radiance."""
    spp = 16
    d, i, _, _ = _gpu_render(scene_dir[name], spp)
    rs = oracle_lib.RefScene(oracle_lib.parse_scene(scene_dir[name]))
    d64, i64 = rs.render_f64(spp, rs.params())
    gpu = ((d + i)[:, :3] / spp).astype(np.float64)
    ref = (d64 + i64)[:, :3] / spp
    close = np.isclose(gpu, ref, rtol=1e-3, atol=1e-4).all(axis=1)
    rmse = float(np.sqrt(((gpu - ref) ** 2).mean()))
    print(f"{name}: {close.mean():.4f} of pixels close, RMSE {rmse:.3e} vs mean {ref.mean():.3e}", flush=True)
    assert close.mean() > 0.97
    assert rmse < 0.05 * ref.mean()


@pytest.mark.parametrize("name", ["cornell_small", "cornell_spheres_128", "terrain_small"])
def test_ray_sort_keeps_results_bit_exact(scene_dir, oracle_lib, name):
    """rt_render_params.ray_sort reorders each bounce's extension rays; every
    path's hits, and so the images and ray counts, stay the oracle's."""
    spp = 4
    rs = oracle_lib.RefScene(oracle_lib.parse_scene(scene_dir[name]))
    s = rt.Scene(scene_dir[name])
    r = rt.Renderer(s, ray_sort=True)
    r.render(spp)
    d = r.read_accum(rt.RT_ACCUM_DIRECT).reshape(-1, 4)
    i = r.read_accum(rt.RT_ACCUM_INDIRECT).reshape(-1, 4)
    rd, ri, rst = rs.render(spp, rs.params())
    assert np.array_equal(_bits(d[:, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[:, :3]), _bits(ri[:, :3]))
    assert r.stats().extension_rays == rst[0] and r.stats().shadow_rays == rst[1]
    assert r.stats().stage_launches[rt.STAGES.index("sort")] == 0  # not profiled: no stage events


def test_ray_sort_counts_graphs_and_batches(scene_dir):
    """Counted traversal totals (each ray walks the same nodes in any order),
    graph replays and pass splits give the unsorted renderer's results."""
    path = scene_dir["cornell_spheres_128"]
    base = rt.Renderer(rt.Scene(path))
    base.render(6, count_traversal=True)
    want = base.stats()
    wd = base.read_accum(rt.RT_ACCUM_DIRECT)
    for kw in ({}, {"graphs": False}, {"batch_spp": 1}, {"batch_spp": 4, "profile": True}):
        r = rt.Renderer(rt.Scene(path), ray_sort=True, **kw)
        r.render(6, count_traversal=True)
        st = r.stats()
        assert (st.extension_node_visits, st.extension_tri_tests, st.shadow_node_visits) == \
            (want.extension_node_visits, want.extension_tri_tests, want.shadow_node_visits), kw
        r2 = rt.Renderer(rt.Scene(path), ray_sort=True, **kw)
        r2.render(6)
        assert np.array_equal(_bits(r2.read_accum(rt.RT_ACCUM_DIRECT)), _bits(wd)), kw
        if kw.get("profile"):
            assert r2.stats().stage_launches[rt.STAGES.index("sort")] > 0


def test_ray_sort_svgf_frames_match_unsorted(scene_dir):
    path = scene_dir["cornell_spheres_rt"]
    a, b = rt.Renderer(rt.Scene(path)), rt.Renderer(rt.Scene(path), ray_sort=True)
    cam = a.camera()
    for f in range(3):
        c = rt.Camera(tuple(cam.target), cam.distance, cam.theta_deg, cam.phi_deg + 1.5 * f)
        a.set_camera(c)
        b.set_camera(c)
        a.render_frame(1)
        b.render_frame(1)
        assert np.array_equal(_bits(a.read_frame(rt.RT_FRAME_DENOISED)), _bits(b.read_frame(rt.RT_FRAME_DENOISED)))


@pytest.mark.parametrize("name", ["cornell_small", "cornell_spheres_128", "terrain_small", "textured_cornell"])
def test_shadow_any_hit_mode_matches_closest_hit(scene_dir, name):
    """The shadow fast mode (rt_render_params.shadow_any_hit) stops occluded
    shadow rays at the first hit clearly in front of the shading point; on
    these scenes every visibility decision, and so every pixel, is the
    closest-hit test's, bit for bit, with fewer node visits."""
    path = scene_dir[name]
    s = rt.Scene(path)
    out = []
    for any_hit in (False, True):
        r = rt.Renderer(s, shadow_any_hit=any_hit)
        r.render(4, count_traversal=True)
        out.append((r.read_accum(rt.RT_ACCUM_DIRECT), r.read_accum(rt.RT_ACCUM_INDIRECT), r.stats()))
    (d0, i0, s0), (d1, i1, s1) = out
    assert np.array_equal(_bits(d0), _bits(d1)) and np.array_equal(_bits(i0), _bits(i1))
    assert s0.shadow_rays == s1.shadow_rays and s0.extension_node_visits == s1.extension_node_visits
    assert s1.shadow_node_visits < s0.shadow_node_visits
    print(f"{name}: shadow node visits {s0.shadow_node_visits} -> {s1.shadow_node_visits}", flush=True)
