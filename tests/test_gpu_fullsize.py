"""Full-size workloads checked on the GPU.

- The large-mesh workload (2.0 M-triangle terrain + subsurface sphere under
  a 2048x1024 HDR sky sampled by alias tables, 1920x1080, 32 spp, depth 16)
  rendered in full; sampled pixels' direct and indirect sums are the oracle's,
  bit for bit.  The oracle renders from the triangles the library imported
  (its own Python OBJ reader is too slow at 2 M triangles; both importers are
  checked against each other and against the reference's tinyobjloader and
  Eigen on smaller files).
- The 64-scene batch (triangles log-uniform 1k-500k, 128^2-1024^2, 1-64 spp)
  through the pool's shared passes; a sample of its jobs, small to large, is
  bit-identical to the same job rendered alone.
- serve(): every result it delivers, batched, plain and plain on resident
  renderers, and with a cap on concurrent real-time jobs, is bit-identical to
  the job rendered alone."""
import json
import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

import cudaraytracer_amd as rt
from cudaraytracer_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def _write_one(job):
    kind, kwargs, out_dir, name = job
    from cudaraytracer_amd import scenes as sc

    return sc.write_scene(getattr(sc, kind)(**kwargs), out_dir, name)


def _write_many(jobs):
    """Write generated scenes in worker processes (spawned: no GPU state)."""
    with ProcessPoolExecutor(max_workers=max(1, min(16, len(jobs), os.cpu_count() or 1)),
                             mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_write_one, jobs))


def test_large_mesh_workload_matches_oracle_on_sampled_pixels(tmp_path, oracle_lib):
    path = _write_many([("terrain_env", {}, str(tmp_path), "terrain_c3")])[0]
    scene = rt.Scene(path)
    info = scene.info()
    assert info.num_triangles == 2005120 and (info.width, info.height, info.spp) == (1920, 1080, 32)
    r = rt.Renderer(scene, env_alias=True)
    r.render(info.spp)
    d = r.read_accum(rt.RT_ACCUM_DIRECT).reshape(-1, 4)
    i = r.read_accum(rt.RT_ACCUM_INDIRECT).reshape(-1, 4)
    assert np.isfinite(d).all() and np.isfinite(i).all() and r.stats().spp_done == 32
    pos, nrm, uv, _ = scene.triangles()
    mesh4, _ = scene.meshes()
    with open(path) as f:
        objects = json.load(f)["objects"]
    geo = oracle_lib.geometry_from_arrays(objects, pos.reshape(-1, 9), nrm.reshape(-1, 9), uv.reshape(-1, 6),
                                          mesh4[:, 2], mesh4[:, 3])
    rs = oracle_lib.RefScene(oracle_lib.parse_scene(path, geo))
    n_bvh, o_bvh = scene.bvh()
    rn, ro = rs.bvh()
    assert np.array_equal(n_bvh, rn) and np.array_equal(o_bvh, ro)  # the GPU-built tree is the oracle's
    pix = np.random.default_rng(31).choice(info.width * info.height, 300, replace=False).astype(np.int32)
    rd, ri, _ = rs.render(info.spp, rs.params(env_alias=True), pixels=pix)
    assert np.array_equal(_bits(d[pix, :3]), _bits(rd[:, :3]))
    assert np.array_equal(_bits(i[pix, :3]), _bits(ri[:, :3]))
    assert float((d + i)[pix, :3].sum()) > 0


def test_batch_mix_jobs_match_alone(tmp_path):
    rng = np.random.default_rng(4)
    specs = []
    for k in range(64):  # the generator's own parameters, as batch_mix draws them
        tris = float(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        width = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.integers(1, 65))
        specs.append(("batch_scene", {"seed": 1000 + k, "target_tris": int(tris), "width": width, "spp": spp},
                      str(tmp_path / f"batch{k:02d}"), f"batch{k:02d}"))
    paths = _write_many(specs)
    sc = [rt.Scene(p) for p in paths]
    pool = rt.Pool(queue_budget_bytes=8 << 30)
    slots = [pool.add_scene(s) for s in sc]
    ids = [pool.submit(rt.Job(slot, spp=s.info().spp)) for slot, s in zip(slots, sc)]
    while pool.step() > 0:
        pass
    infos = [s.info() for s in sc]
    order = sorted(range(64), key=lambda k: infos[k].width * infos[k].width * infos[k].spp)
    sample = sorted(set(order[::9] + [order[-1]]))
    for k in sample:
        inf = infos[k]
        st = pool.status(ids[k])
        assert st.done == 1 and st.spp_done == inf.spp
        r = rt.Renderer(sc[k])
        r.render(inf.spp)
        for which in (rt.RT_ACCUM_DIRECT, rt.RT_ACCUM_INDIRECT):
            got = pool.read_accum(ids[k], which, inf.width, inf.height)
            assert np.array_equal(_bits(got), _bits(r.read_accum(which))), (k, which)
        r.close()
    print(f"checked jobs {sample} of 64", flush=True)


def _cams(base, n, phi0=0.0):
    return [rt.Camera(tuple(base.target), base.distance, base.theta_deg, base.phi_deg + phi0 + 1.5 * f)
            for f in range(n)]


@pytest.mark.parametrize("mode,cap", [(True, 0), (True, 2), (False, 0), ("resident", 0)])
def test_serve_results_match_alone(scene_dir, mode, cap):
    names = ["cornell_spheres_rt", "terrain_small", "cornell_small"]
    sc = [rt.Scene(scene_dir[n]) for n in names]
    pool = rt.Pool(max_realtime_jobs=cap)
    slots = [pool.add_scene(s) for s in sc]
    base = [rt.Renderer(s).camera() for s in sc]
    reqs, want, res = [], [], []
    for k in range(8):
        s = k % 3
        info = sc[s].info()
        if s == 2:
            job = rt.Job(slots[s], spp=3 + k)
            r = rt.Renderer(sc[s])
            r.render(3 + k)
            want.append(r.read_accum(rt.RT_ACCUM_MEAN))
        else:
            cams = _cams(base[s], 3, phi0=2.0 * k)
            job = rt.Job(slots[s], spp=1, frames=3, cameras=cams)
            r = rt.Renderer(sc[s])
            for c in cams:
                r.set_camera(c)
                r.render_frame(1)
            want.append(r.read_frame(rt.RT_FRAME_DENOISED))
        r.close()
        reqs.append((0.003 * k, job))
        res.append(np.full((info.height, info.width, 4), np.nan, np.float32))
    rep, _, _ = pool.serve(reqs, batched=mode, results=res)
    assert rep["jobs"] == 8
    for k in range(8):
        assert np.array_equal(_bits(res[k]), _bits(want[k].reshape(res[k].shape))), (mode, cap, k)
