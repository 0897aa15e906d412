"""The multi-job pool (batching and serving) against the single-scene renderer.

A job's results must not depend on what else shares its passes: every
accumulate job's sums and every real-time job's frame buffers are required to
be bit-identical to the same job rendered alone by the single-scene renderer
(itself bit-exact against the oracle), including when the pool's small path
capacity splits a job's samples over several passes."""
import numpy as np
import pytest

import cudaraytracer_amd as rt

pytestmark = pytest.mark.gpu

ACCUM = [("cornell_small", 3), ("cornell_spheres_128", 4), ("terrain_small", 5), ("cornell_spheres_rt", 7),
         ("textured_cornell", 3)]


def _bits_equal(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        ok = np.array_equal(a.view(np.int32), b.view(np.int32))
    else:
        ok = np.array_equal(a, b)
    assert ok, f"{what}: {np.count_nonzero(a != b)} of {a.size} differ"


def _alone_accum(scene, spp):
    r = rt.Renderer(scene)
    r.render(spp)
    return r.read_accum(rt.RT_ACCUM_DIRECT), r.read_accum(rt.RT_ACCUM_INDIRECT), r.stats()


@pytest.mark.parametrize("queue_mb", [34, 4096])
def test_pool_accumulate_jobs_match_alone(scene_dir, queue_mb):
    """34 MB of queues = ~150 k paths: jobs are split over many passes and
    share them; 4 GB: every job fits one pass."""
    scenes = {n: rt.Scene(scene_dir[n]) for n, _ in ACCUM}
    pool = rt.Pool(queue_budget_bytes=queue_mb << 20)
    slots = {n: pool.add_scene(s) for n, s in scenes.items()}
    jobs = {n: pool.submit(rt.Job(slots[n], spp=spp)) for n, spp in ACCUM}
    passes = 0
    while pool.step() > 0:
        passes += 1
    print(f"queue {queue_mb} MB: {passes} passes", flush=True)
    ext = 0
    for n, spp in ACCUM:
        info = scenes[n].info()
        st = pool.status(jobs[n])
        assert st.done == 1 and st.spp_done == spp
        d_alone, i_alone, s_alone = _alone_accum(scenes[n], spp)
        _bits_equal(pool.read_accum(jobs[n], rt.RT_ACCUM_DIRECT, info.width, info.height), d_alone, f"{n} direct")
        _bits_equal(pool.read_accum(jobs[n], rt.RT_ACCUM_INDIRECT, info.width, info.height), i_alone, f"{n} indirect")
        ext += s_alone.extension_rays
    assert pool.stats().extension_rays == ext
    if queue_mb == 34:
        assert passes >= 3  # 405 k paths in passes of ~156 k: split and shared


def _path(cam, n, dphi=1.5, dtheta=-0.5):
    return [rt.Camera(tuple(cam.target), cam.distance, cam.theta_deg + dtheta * k, cam.phi_deg + dphi * k)
            for k in range(n)]


FRAME_BUFFERS = [rt.RT_FRAME_DENOISED, rt.RT_FRAME_HISTORY_DIRECT, rt.RT_FRAME_HISTORY_INDIRECT,
                 rt.RT_FRAME_MOMENTS_DIRECT, rt.RT_FRAME_HISTORY_LENGTH, rt.RT_FRAME_VARIANCE_INDIRECT,
                 rt.RT_FRAME_DIRECT, rt.RT_FRAME_INDIRECT, rt.RT_FRAME_G_MOTION_DEPTH, rt.RT_FRAME_G_NORMAL_MAT]


@pytest.mark.parametrize("pipelined", [True, False])
def test_pool_realtime_jobs_match_alone(scene_dir, pipelined):
    names = ["cornell_spheres_rt", "terrain_small"]
    nframes = 4
    scenes = {n: rt.Scene(scene_dir[n]) for n in names}
    # Alone: the single-scene renderer's frame loop, every frame's buffers kept.
    alone, paths = {}, {}
    for n in names:
        r = rt.Renderer(scenes[n])
        paths[n] = _path(r.camera(), nframes)
        frames = []
        for cam in paths[n]:
            r.set_camera(cam)
            r.render_frame(1)
            frames.append({w: r.read_frame(w) for w in FRAME_BUFFERS})
        alone[n] = frames
    # In the mix: both real-time jobs plus an accumulate job share the passes.
    pool = rt.Pool(pipelined=pipelined)
    slots = {n: pool.add_scene(scenes[n]) for n in names}
    extra = rt.Scene(scene_dir["cornell_small"])
    jobs = {n: pool.submit(rt.Job(slots[n], spp=1, frames=nframes, cameras=paths[n])) for n in names}
    acc_job = pool.submit(rt.Job(pool.add_scene(extra), spp=6))
    seen = {n: 0 for n in names}
    for _ in range(4 * nframes + 4):
        assert pool.step() > 0
        for n in names:
            info = scenes[n].info()
            done = pool.status(jobs[n]).frames_done
            assert done <= seen[n] + 1
            if done == seen[n] + 1:  # a frame completed in this tick: check it
                f = seen[n]
                for w in FRAME_BUFFERS:
                    got = pool.read_frame(jobs[n], w, info.width, info.height)
                    _bits_equal(got, alone[n][f][w], f"{n} frame {f} buffer {w}")
                seen[n] = done
        if all(v == nframes for v in seen.values()):
            break
    assert all(v == nframes for v in seen.values())
    pool.run()
    assert pool.status(acc_job).done == 1
    d_alone, _, _ = _alone_accum(extra, 6)
    _bits_equal(pool.read_accum(acc_job, rt.RT_ACCUM_DIRECT, 256, 256), d_alone, "accumulate job in the mix")


def test_pool_admission_and_errors(scene_dir):
    scene = rt.Scene(scene_dir["cornell_spheres_rt"])
    pool = rt.Pool(hbm_budget_bytes=64 << 20)
    slot = pool.add_scene(scene)
    cam = rt.Renderer(scene).camera()
    frames = _path(cam, 2)
    ids = []
    with pytest.raises(rt.RtError) as e:
        for _ in range(200):  # 96x64 real-time jobs take ~2.2 MB each
            ids.append(pool.submit(rt.Job(slot, spp=1, frames=2, cameras=frames)))
    assert e.value.code == rt.RT_ERR_BUSY
    assert len(ids) > 5
    pool.release(ids[0])
    ids[0] = pool.submit(rt.Job(slot, spp=1, frames=2, cameras=frames))
    with pytest.raises(rt.RtError) as e:
        pool.submit(rt.Job(slot + 7, spp=1))
    assert e.value.code == -1
    with pytest.raises(rt.RtError):
        pool.submit(rt.Job(slot, spp=1, frames=2))  # no cameras
    pool.run()
    assert all(pool.status(i).done == 1 for i in ids)


def test_pool_serve_batched_and_plain(scene_dir):
    names = ["cornell_spheres_rt", "terrain_small", "cornell_small"]
    scenes = [rt.Scene(scene_dir[n]) for n in names]
    pool = rt.Pool()
    slots = [pool.add_scene(s) for s in scenes]
    reqs = []
    for k in range(9):
        s = k % 3
        if s == 2:
            reqs.append((0.004 * k, rt.Job(slots[s], spp=4)))
        else:
            cam = rt.Renderer(scenes[s]).camera()
            reqs.append((0.004 * k, rt.Job(slots[s], spp=1, frames=3, cameras=_path(cam, 3))))
    for batched in (True, False):
        rep, frame_ms, job_ms = pool.serve(reqs, batched=batched)
        print("batched" if batched else "plain", rep, flush=True)
        assert rep["jobs"] == 9 and rep["frames"] == 6 * 3 + 3
        assert rep["makespan_s"] > 0 and rep["jobs_per_s"] > 0
        assert all(len(f) == (3 if k % 3 != 2 else 1) for k, f in enumerate(frame_ms))
        assert (job_ms > 0).all()
        for f, j in zip(frame_ms, job_ms):
            assert abs(f.sum() - j) < 1e-6 * max(1.0, j)


def test_pool_tuning_keeps_results_and_autotune_runs(scene_dir):
    """Launch shape, pass size, tile shape and the ray sort change speed only: results stay
    bit-identical; the autotuner picks a valid tuning from a sample."""
    scene = rt.Scene(scene_dir["cornell_spheres_rt"])
    r = rt.Renderer(scene)
    frames = _path(r.camera(), 3)
    alone = []
    for cam in frames:
        r.set_camera(cam)
        r.render_frame(1)
        alone.append(r.read_frame(rt.RT_FRAME_DENOISED))
    d_alone, _, _ = _alone_accum(scene, 5)
    info = scene.info()
    pool = rt.Pool()
    slot = pool.add_scene(scene)
    for scale, paths, tw, sort in ((0.5, 0, 8, False), (1.25, 20000, 64, True), (1.0, 0, 32, True)):
        pool.set_tuning(scale, paths, tw, sort)
        t = pool.tuning()
        assert (t.grid_scale, t.pass_paths, t.tile_width, t.ray_sort) == (scale, paths, tw, int(sort))
        acc = pool.submit(rt.Job(slot, spp=5))
        rtj = pool.submit(rt.Job(slot, spp=1, frames=3, cameras=frames))
        f = 0
        while pool.step() > 0:
            st = pool.status(rtj)
            if st.frames_done > f:
                _bits_equal(pool.read_frame(rtj, rt.RT_FRAME_DENOISED, info.width, info.height), alone[f],
                            f"tuning {scale, paths, tw, sort} frame {f}")
                f = st.frames_done
        assert f == 3
        _bits_equal(pool.read_accum(acc, rt.RT_ACCUM_DIRECT, info.width, info.height), d_alone, "accumulate")
        pool.release(acc)
        pool.release(rtj)
    rep = pool.autotune([rt.Job(slot, spp=4), rt.Job(slot, spp=1, frames=2, cameras=frames[:2])])
    print("autotune", rep, flush=True)
    assert rep["runs"] >= 4 and rep["best_ms"] <= rep["baseline_ms"] + 1e-9
    assert rep["best"]["tile_width"] in (8, 16, 32, 64) and rep["best"]["ray_sort"] in (0, 1)
