"""Benchmark: wavefront path tracing throughput on one MI355X.

Headline workload (one step): the large-mesh scene -- a 2,000,000-triangle fBm
terrain plus a random-walk subsurface icosphere under a procedural 2048x1024
HDR sky (peak 1e4 sun lobe) sampled through alias tables -- at 1920x1080, 32
samples per pixel, max depth 16, rendered in full.  The same render with the
reference's Burley subsurface profile is timed beside it (sss_burley).  Inputs (BVH, triangles, env map,
queues) are resident in HBM before timing starts.

Prints one JSON line.  `value` is Mrays/s over the timed steps (extension +
shadow rays / GPU wall time).  `roofline` is for the kernel with the largest
share of GPU time, from per-launch HIP events recorded on the renderer's own
stream and the algorithmic-bytes model below (counted node visits and triangle
tests times their record sizes); `traffic` is the HBM bytes per launch that
rocprofv3's FETCH_SIZE / WRITE_SIZE counters measured for the same kernel,
scaled by what they read for a plain 1 GiB copy (the tables under profiles/,
written by --traffic-from-pmc; the shade stage is counted per bounce, all
material classes together).  `cpu_baseline` times the CPU oracle (OpenMP, all
host cores) on a bounded sample of the same workload.  Secondary sections: the
spheres workload (512x512, 64 spp), the small Cornell box on the GPU and the
CPU, the real-time SVGF frame loop, the 64-scene batch and the serving stream.

Other modes: --only NAME runs one secondary workload (frames, c1, c2, batched,
serving, bvh: the 2 M-triangle SAH build on the host and the GPU, textured: the
cost of diffuse textures) and prints its JSON; --calibrate-copy runs the 1 GiB
copy that scales the counters; --traffic-from-pmc turns four rocprofv3 --pmc
databases into the traffic table.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes: the records each stage must move, by unit.
#   extend: the ray (origin + direction, 32 B) per extension ray; the hit record
#     (16 B) written per hit; per miss under an environment light, the path
#     weight (16 B), the four RGB texels of the bilinear lookup (48 B) and the
#     read-modify-write of the path's sum (32 B); 64 B per interior node whose
#     child pair is tested; 36 B of vertex data per triangle test.
#   shade (per hit): the class-list entry (4 B), the ray and weight (48 B), the hit
#     record (16 B) and the hit triangle's 128 B attribute line; the next bounce's
#     ray and weight (48 B) per survivor; the 64 B shadow record per shadow ray;
#     the read-modify-write of the path's sum (32 B) per light hit; per
#     next-event estimate that samples the environment, the two alias-table
#     entries (16 B), the four pdf-table floats (16 B) and the bilinear texels (48 B).
#   shadow: the 64 B record per shadow ray, the nodes and triangles it tests,
#     and the read-modify-write of the path's sum (32 B) per visible light point.
EXT_RAY_BYTES, HIT_BYTES, NODE_BYTES, TRI_BYTES = 32, 16, 64, 36
MISS_ENV_BYTES = 16 + 48 + 32
SHADE_HIT_BYTES, ATTR_LINE_BYTES, SURVIVOR_BYTES, SHADOW_REC_BYTES, SUM_RMW_BYTES = 68, 128, 48, 64, 32
ENV_NEE_BYTES = 16 + 16 + 48
# With the ray sort on, after the first bounce extend also reads a 4 B permutation
# entry and writes a 32 B copy of the ray at its sorted position (hits only), shade
# reads the permutation entry (for the weight) and writes a 2 B sort key per survivor.
PERM_BYTES, RAY_COPY_BYTES, SORT_KEY_BYTES = 4, 32, 2


TRAFFIC_STAGES = ("k_extend", "k_shade", "k_shadow")
C2_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_cornell_spheres_512_64spp.txt")
C3_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_terrain_2M_1080p_32spp.txt")


def _stage_of(kernel: str):
    """Stage of a measured kernel instance: only the plain (uncounted,
    single-scene) instantiations the timed renders launch.  The extend and
    shadow stages are their kernels (with or without straggler suspension)
    and the continuation launches that finish the suspended searches."""
    import re

    if re.match(r"(?:rt::)?(?:k_extend<false, false(?:, \d+)?|k_extend_cont<false, false)>$", kernel):
        return "k_extend"
    if re.match(r"(?:rt::)?(?:k_shadow<false, false(?:, (?:false|0)(?:, \d+)?)?|k_shadow_cont<false, false)>$", kernel):
        return "k_shadow"
    if re.match(r"(?:rt::)?k_shade<(?:\d, )?false>$", kernel):
        return "k_shade"
    return None


def _kernel_name(name: str) -> str:
    import re

    return re.sub(r"\(.*", "", name.replace("(anonymous namespace)::", "")).replace("void ", "").strip()


def kernel_table(db, top: int = 30) -> str:
    """Markdown summary of a rocprofv3 --kernel-trace database: calls, total and
    average duration, share of GPU time and launch resources per kernel."""
    import sqlite3

    con = sqlite3.connect(db)
    res = {kid: (vg, ag, sg) for kid, vg, ag, sg in
           con.execute("select id, arch_vgpr_count, accum_vgpr_count, sgpr_count from rocpd_info_kernel_symbol")}
    agg, total = {}, 0.0
    for name, dur, grid, wg, lds, scr, kid in con.execute(
            "select name, duration, grid_x*grid_y*grid_z, workgroup_x*workgroup_y*workgroup_z, lds_size, "
            "scratch_size, kernel_id from kernels"):
        a = agg.setdefault(_kernel_name(name), [0, 0.0, grid, wg, lds, scr, res.get(kid, ("", "", ""))])
        a[0] += 1
        a[1] += dur
        total += dur
    lines = ["| kernel | calls | total us | avg us | % GPU time | grid | block | LDS B | VGPR | AGPR | SGPR | scratch |",
             "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for n, (cnt, d, grid, wg, lds, scr, (vg, ag, sg)) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        lines.append(f"| {n} | {cnt} | {d / 1e3:.1f} | {d / 1e3 / cnt:.1f} | {100 * d / total:.2f} | {grid} | {wg} | "
                     f"{lds} | {vg} | {ag} | {sg} | {scr} |")
    return "\n".join(lines)


def counter_table(kind: str, db, top: int = 12) -> str:
    """Markdown summary of a rocprofv3 --pmc database, per kernel.  kind "sq": SQ
    instruction mix per wave; kind "l2": L2 hit rate and VALU lane utilisation."""
    import sqlite3
    from collections import defaultdict

    con = sqlite3.connect(db)
    agg, disp, dur = defaultdict(lambda: defaultdict(float)), defaultdict(set), {}
    for kn, d, cn, v, du in con.execute(
            "select kernel_name, dispatch_id, counter_name, value, duration from counters_collection"):
        n = _kernel_name(kn)
        agg[n][cn] += v
        disp[n].add(d)
        dur[(n, d)] = du
    tot = defaultdict(float)
    for (n, _), du in dur.items():
        tot[n] += du
    rows = sorted(agg, key=lambda k: -tot[k])[:top]
    if kind == "l2":
        lines = ["| kernel | dispatches | L2 hit | VALU lane utilisation |", "|---|---|---|---|"]
        for n in rows:
            a = agg[n]
            hit, miss = a.get("TCC_HIT_sum", 0), a.get("TCC_MISS_sum", 0)
            util = a.get("SQ_THREAD_CYCLES_VALU", 0) / max(64 * a.get("SQ_ACTIVE_INST_VALU", 0), 1)
            lines.append(f"| {n} | {len(disp[n])} | {100 * hit / max(hit + miss, 1):.1f} % | {100 * util:.1f} % |")
        return "\n".join(lines)
    lines = ["| kernel | dispatches | waves | VALU / wave | VMEM rd / wave | LDS / wave | wait / wave cycles | "
             "VALU wave-instr / us |", "|---|---|---|---|---|---|---|---|"]
    for n in rows:
        a = agg[n]
        w = a.get("SQ_WAVES", 0) or 1
        lines.append(f"| {n} | {len(disp[n])} | {int(w)} | {a.get('SQ_INSTS_VALU', 0) / w:.0f} | "
                     f"{a.get('SQ_INSTS_VMEM_RD', 0) / w:.0f} | {a.get('SQ_INSTS_LDS', 0) / w:.0f} | "
                     f"{a.get('SQ_WAIT_INST_ANY', 0) / max(a.get('SQ_WAVE_CYCLES', 0), 1):.3f} | "
                     f"{a.get('SQ_INSTS_VALU', 0) / (tot[n] / 1e3):.0f} |")
    return "\n".join(lines)


def traffic_from_pmc(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_bytes=1 << 30):
    """HBM bytes per launch of each stage from rocprofv3 --pmc databases.
    FETCH_SIZE and WRITE_SIZE are KiB per dispatch; each is scaled by the
    ratio of the true size of a plain device copy to what the same counter
    read for it."""
    import re
    import sqlite3

    def per_kernel(db):
        con = sqlite3.connect(db)
        out = {}
        for name, disp, val in con.execute("select kernel_name, dispatch_id, value from counters_collection"):
            k = _kernel_name(name)
            d = out.setdefault(k, {})
            d[disp] = d.get(disp, 0.0) + val
        return out

    def copy_kib(db):
        rows = [sum(v.values()) / len(v) for k, v in per_kernel(db).items() if "copyBuffer" in k]
        if not rows:
            raise SystemExit(f"{db}: no copy kernel found")
        return max(rows)

    f_scale = copy_bytes / (copy_kib(copy_fetch_db) * 1024)
    w_scale = copy_bytes / (copy_kib(copy_write_db) * 1024)
    fetch, write = per_kernel(fetch_db), per_kernel(write_db)
    lines = [f"fetch_scale {f_scale:.4f}", f"write_scale {w_scale:.4f}", "",
             "k_shade is per bounce: its class kernels' bytes summed, divided by the bounces (the single-scene",
             "k_extend dispatches, counting renders included, as their shade launches are the same kernels);",
             "k_extend and k_shadow are per bounce too: each launch and its continuation launch summed", "",
             "stage      dispatches  fetch_B_per_launch  write_B_per_launch  hbm_B_per_launch"]
    n_bounces = sum(len(fetch[k]) for k in fetch if re.match(r"(?:rt::)?k_extend<(?:true|false), false(?:, \d+)?>$", k))
    n_ext = sum(len(fetch[k]) for k in fetch if re.match(r"(?:rt::)?k_extend<false, false(?:, \d+)?>$", k))
    n_sh = sum(len(fetch[k]) for k in fetch if _stage_of(k) == "k_shadow" and "_cont" not in k)
    for st in TRAFFIC_STAGES:
        fk = [k for k in fetch if _stage_of(k) == st]
        wk = [k for k in write if _stage_of(k) == st]
        nf = sum(len(fetch[k]) for k in fk)
        nw = sum(len(write[k]) for k in wk)
        if not nf or not nw:
            continue
        per = (n_bounces if st == "k_shade" and n_bounces else
               n_ext if st == "k_extend" and n_ext else
               n_sh if st == "k_shadow" and n_sh else None)
        fb = sum(sum(fetch[k].values()) for k in fk) * 1024 * f_scale / (per or nf)
        wb = sum(sum(write[k].values()) for k in wk) * 1024 * w_scale / (per or nw)
        lines.append(f"{st:10s} {per or nf:11d} {fb:19.0f} {wb:19.0f} {fb + wb:17.0f}")
    def inst(k):
        f, w = fetch.get(k, {}), write.get(k, {})
        fb = sum(f.values()) * 1024 * f_scale / max(1, len(f))
        wb = sum(w.values()) * 1024 * w_scale / max(1, len(w))
        return f"  {k:36s} {len(f):5d} {fb:15.0f} {wb:15.0f}"

    names = sorted(set(fetch) | set(write))
    lines += ["", "instance-level (scaled bytes per dispatch):"]
    lines += [inst(k) for k in names if _stage_of(k) is not None]
    others = sorted((k for k in names if _stage_of(k) is None),
                    key=lambda k: -(sum(fetch.get(k, {}).values()) + sum(write.get(k, {}).values())))
    lines += ["", "other kernels, by total bytes (scaled bytes per dispatch):"]
    lines += [inst(k) for k in others[:12]]
    return "\n".join(lines)


def load_traffic(path):
    """Stage -> HBM bytes per launch from a traffic table, {} if absent."""
    out = {}
    if not path or not os.path.exists(path):
        return out
    for line in open(path):
        tok = line.split()
        if len(tok) == 5 and tok[0] in TRAFFIC_STAGES and tok[1].isdigit():
            out[tok[0]] = int(float(tok[4]))
    return out


def calibrate_copy(nbytes: int = 1 << 30):
    """A plain device-to-device copy of nbytes, three times: the counter scale."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))


def _stage_roofline(name, bytes_per_launch, ms_per_launch, traffic=None):
    achieved = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": int(bytes_per_launch), "ms_per_launch": round(ms_per_launch, 5)}


# SVGF, per pixel, each pixel's data read once: temporal reads this frame's sums
# (32), last frame's history, moments and length (52) and both G-buffers
# (2 x 48) and writes history, moments and length (52), plus the variance
# pass's output (32) for a pixel whose history reached 4 frames; the variance
# pass estimates only the other pixels, each reading its list entry (4), colour
# (32), moments (16), length (4), filter guide (24) and luminance guide (16) and
# writing 32; one a-trous pass reads colour (32), length (4) and G-buffer (32)
# and writes 32.  The G-buffer kernel writes 48 B per pixel (its traversal is
# not counted).
TEMPORAL_PX_BYTES, TEMPORAL_COPY_BYTES, VARIANCE_PX_BYTES, ATROUS_PX_BYTES, GBUFFER_PX_BYTES = 232, 32, 128, 100, 48


def frames_bench(frames: int = 10, width: int = 1280, height: int = 720, env_alias: bool = True):
    """Real-time loop on the terrain + subsurface + HDR environment scene
    (2 M triangles): `frames` frames of an orbiting camera at 1 spp with SVGF
    (temporal, variance, 5 a-trous passes), the environment sampled through
    the alias tables.  Returns per-stage ms per frame."""
    import cudaraytracer_amd as rt
    from cudaraytracer_amd import scenes

    tmp = tempfile.mkdtemp(prefix="rtframes_")
    path = scenes.write_scene(scenes.terrain_env(width=width, height=height, spp=1), tmp, "terrain_frames")
    scene = rt.Scene(path)
    info = scene.info()

    # Wall-clock frame latency (synchronised per frame), then a profiled pass for stage times.
    r = rt.Renderer(scene, env_alias=env_alias)
    base = r.camera()
    r.render_frame(1)
    r.synchronize()
    lat = []
    for k in range(1, frames + 1):
        r.set_camera(rt.Camera(tuple(base.target), base.distance, base.theta_deg, base.phi_deg + 0.5 * k))
        t0 = time.perf_counter()
        r.render_frame(1)
        r.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    den = r.read_frame(rt.RT_FRAME_DENOISED)
    assert np.isfinite(den).all()
    rp = rt.Renderer(scene, profile=True, env_alias=env_alias)
    rp.render_frame(1)
    rp.synchronize()
    rp.reset_stats()
    short = 0  # pixels the variance pass estimated (history < 4 frames), over the frames
    for k in range(1, frames + 1):
        rp.set_camera(rt.Camera(tuple(base.target), base.distance, base.theta_deg, base.phi_deg + 0.5 * k))
        rp.render_frame(1)
        short += int((rp.read_frame(rt.RT_FRAME_HISTORY_LENGTH) < 4).sum())
    st = rp.stats()
    ms = {name: st.stage_ms[i] / frames for i, name in enumerate(rt.STAGES)}
    # Held camera: after 4 frames every history is long unless the view is disoccluded, and
    # the variance pass visits only the tiles that still hold a short history.
    for _ in range(4):
        rp.render_frame(1)
    rp.synchronize()
    rp.reset_stats()
    held = 5
    held_short = 0
    for _ in range(held):
        rp.render_frame(1)
        held_short += int((rp.read_frame(rt.RT_FRAME_HISTORY_LENGTH) < 4).sum())
    sh = rp.stats()
    px = width * height
    short_pf = short / frames
    atrous_launch_ms = st.stage_ms[8] / frames / 5
    ach = ATROUS_PX_BYTES * px / (atrous_launch_ms * 1e-3) / 1e9
    per_stage_gbs = {
        "temporal": (TEMPORAL_PX_BYTES * px + TEMPORAL_COPY_BYTES * (px - short_pf)) / (ms["temporal"] * 1e-3) / 1e9,
        "variance": VARIANCE_PX_BYTES * short_pf / (ms["variance"] * 1e-3) / 1e9,
        "atrous": ach,
    }
    return {
        "workload": f"terrain_env {info.num_triangles} tris + SSS sphere + 2048x1024 env, {width}x{height}, "
                    f"{frames}-frame orbit, 1 spp/frame, SVGF 5 a-trous passes, env sampling: "
                    f"{'alias tables' if env_alias else 'CDF search'}",
        "frame_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "frame_ms_max": round(float(np.max(lat)), 3),
        "stage_ms_per_frame": {k: round(v, 4) for k, v in ms.items()},
        "svgf_gbs": {k: round(v, 1) for k, v in per_stage_gbs.items()},
        "variance_pixels_per_frame": round(short_pf),
        "held_camera": {"variance_ms_per_frame": round(sh.stage_ms[7] / held, 4),
                        "temporal_ms_per_frame": round(sh.stage_ms[6] / held, 4),
                        "variance_pixels_per_frame": round(held_short / held)},
        "atrous_roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(ach / HBM_PEAK_GBS, 4)},
    }


def stage_bytes(sc, scene, primary: int, sort: bool) -> dict:
    """Algorithmic bytes of one counted render per stage, from its path counts
    (rt_stats_out of a count_traversal render)."""
    info = scene.info()
    env = scene.envmap().shape[0] > 0
    E, M = sc.extension_rays, sc.miss_paths
    hits = E - M
    cls = sc.as_dict()["class_paths"]
    nee_paths = cls["diffuse"] + cls["ggx"] + cls["subsurface"]
    env_nee = nee_paths / max(1, info.num_lights) if env else 0.0
    survivors = E - primary
    sorted_hits = (survivors * hits / max(1, E)) if sort else 0.0
    ext = (E * EXT_RAY_BYTES + hits * HIT_BYTES + (M * MISS_ENV_BYTES if env else 0)
           + sc.extension_node_visits * NODE_BYTES + sc.extension_tri_tests * TRI_BYTES)
    ext += (survivors * PERM_BYTES + sorted_hits * RAY_COPY_BYTES) if sort else 0
    shade = (hits * (SHADE_HIT_BYTES + ATTR_LINE_BYTES) + survivors * SURVIVOR_BYTES
             + sc.shadow_rays * SHADOW_REC_BYTES + cls["light"] * SUM_RMW_BYTES + env_nee * ENV_NEE_BYTES)
    shade += (sorted_hits * PERM_BYTES + survivors * SORT_KEY_BYTES) if sort else 0
    shadow = (sc.shadow_rays * SHADOW_REC_BYTES + sc.shadow_node_visits * NODE_BYTES
              + sc.shadow_tri_tests * TRI_BYTES + sc.shadow_lit * SUM_RMW_BYTES)
    return {"k_extend": ext, "k_shade": shade, "k_shadow": shadow}


def measure_render(scene, steps: int, warmup: int, traffic_file=None, **renderer_kw):
    """Full renders of a loaded scene at its own spp: `warmup` untimed, then
    `steps` timed (synchronised on both sides); then one profiled step for the
    per-stage HIP-event times and one counted step for the algorithmic bytes.
    Returns throughput, per-stage rooflines and traversal statistics."""
    import cudaraytracer_amd as rt

    info = scene.info()
    spp = info.spp
    sort = bool(renderer_kw.get("ray_sort", False))
    r = rt.Renderer(scene, **renderer_kw)
    for _ in range(warmup):
        r.reset()
        r.render(spp)
    r.synchronize()
    r.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.reset()
        r.render(spp)
    r.synchronize()
    elapsed = time.perf_counter() - t0
    st = r.stats()
    rays = st.extension_rays + st.shadow_rays
    img = r.read_accum(rt.RT_ACCUM_MEAN)
    assert np.isfinite(img).all()
    rp = rt.Renderer(scene, profile=True, **renderer_kw)
    rp.render(spp)
    rp.synchronize()
    rp.reset()
    rp.reset_stats()
    prof_steps = max(1, min(3, steps))
    for _ in range(prof_steps):
        rp.reset()
        rp.render(spp)
    sp = rp.stats()
    rc = rt.Renderer(scene, **renderer_kw)
    rc.render(spp, count_traversal=True)
    sc = rc.stats()
    traffic = load_traffic(traffic_file)
    model = stage_bytes(sc, scene, info.width * info.height * spp, sort)
    ext_bytes, shade_bytes, sh_bytes = model["k_extend"], model["k_shade"], model["k_shadow"]
    stages = {}
    bounces = sp.stage_launches[1] / prof_steps  # one extend launch per bounce
    for name, nbytes, idx in (("k_extend", ext_bytes, 1), ("k_shade", shade_bytes, 2), ("k_shadow", sh_bytes, 3)):
        # shade: one "launch" = one bounce, every material class together
        launches = bounces if name == "k_shade" else sp.stage_launches[idx] / prof_steps
        ms = sp.stage_ms[idx] / prof_steps
        stages[name] = _stage_roofline(name, nbytes / launches, ms / launches, traffic.get(name))
        stages[name]["ms_per_step"] = round(ms, 3)
    return {
        "mrays_s": round(rays / elapsed / 1e6, 2),
        "ms_per_step": round(elapsed * 1e3 / steps, 3),
        "rays_per_step": int(rays // steps),
        "batch_spp": st.batch_spp,
        "path_counts": {"extension_rays": sc.extension_rays, "misses": sc.miss_paths,
                        "shaded_hits": dict(sc.as_dict()["class_paths"]), "shadow_rays": sc.shadow_rays,
                        "shadow_lit": sc.shadow_lit, "subsurface_rays": sc.probe_rays,
                        "bounces": int(round(bounces))},
        "nodes_per_ext_ray": round(sc.extension_node_visits / max(1, sc.extension_rays), 2),
        "tris_per_ext_ray": round(sc.extension_tri_tests / max(1, sc.extension_rays), 2),
        "stage_ms_per_step": {n: round(sp.stage_ms[i] / prof_steps, 3) for i, n in enumerate(rt.STAGES)
                              if sp.stage_launches[i]},
        "stages": stages,
        "dominant": max(stages.values(), key=lambda x: x["ms_per_step"]),
    }


def c3_scene_path():
    tmp = tempfile.mkdtemp(prefix="rtc3_")
    return _generate([("terrain_env", {}, tmp, "terrain_c3")])[0]


def burley_variant(path: str) -> str:
    """A copy of a scene JSON (same directory, same OBJ files) whose
    random-walk subsurface materials use the Burley profile instead."""
    with open(path) as f:
        doc = json.load(f)
    for m in doc["materials"]:
        if m.get("subsurface_method") == "random_walk":
            m["subsurface_method"] = "burley"
    out = os.path.splitext(path)[0] + "_burley.json"
    with open(out, "w") as f:
        json.dump(doc, f)
    return out


def c2_bench(steps: int = 5, warmup: int = 1, ray_sort: bool = True):
    """The spheres workload: Cornell shell + 4 icospheres (20,516 triangles, GGX,
    IOR-1.5 dielectric, MIS), 512x512, 64 spp, depth 12, the ray sort on."""
    import cudaraytracer_amd as rt
    from cudaraytracer_amd import scenes

    path = scenes.write_scene(scenes.cornell_spheres(), tempfile.mkdtemp(prefix="rtc2_"), "cornell_spheres")
    scene = rt.Scene(path)
    m = measure_render(scene, steps, warmup, C2_TRAFFIC, ray_sort=ray_sort)
    m["workload"] = ("cornell_spheres: Cornell shell + 4 icospheres (subdiv 4), 20516 tris, GGX + dielectric + MIS, "
                     f"512x512, 64 spp, depth 12, RR after depth 3, ray sort {'on' if ray_sort else 'off'}")
    m["roofline"] = {k: m["dominant"][k] for k in ("kernel", "bound", "achieved", "peak", "unit", "frac", "traffic")}
    del m["dominant"]
    return m, path


def c1_bench(steps: int = 20, warmup: int = 2, cpu: bool = True):
    """The small Cornell box (36 triangles, 256x256, 16 spp, depth 8, roulette
    after depth 3) on the GPU and, in full, on the CPU oracle."""
    import cudaraytracer_amd as rt
    from cudaraytracer_amd import scenes

    path = scenes.write_scene(scenes.cornell_small(), tempfile.mkdtemp(prefix="rtc1_"), "cornell_small")
    scene = rt.Scene(path)
    m = measure_render(scene, steps, warmup)
    out = {"workload": "cornell_small: 5 wall quads + short and tall box + light quad, 36 tris, 256x256, 16 spp, "
                       "depth 8, RR after depth 3",
           "gpu_mrays_s": m["mrays_s"], "gpu_ms_per_render": m["ms_per_step"], "rays_per_render": m["rays_per_step"]}
    if cpu:
        out["cpu"] = cpu_baseline(path, scene.info().spp, full=True)
        out["gpu_over_cpu"] = round(out["gpu_mrays_s"] / out["cpu"]["value"], 1)
    return out


def c3_extra(path):
    import cudaraytracer_amd as rt

    scene = rt.Scene(path)
    return {"bvh_build": bvh_build_bench(path, gpu_scene=scene)}


def bvh_build_bench(path, gpu_scene=None):
    """SAH BVH build of a loaded scene on the host and on the GPU: wall time of
    each (the GPU one warm, after a first build in the process), its phases,
    and whether the two trees are identical."""
    import cudaraytracer_amd as rt

    if gpu_scene is None:
        gpu_scene = rt.Scene(path, bvh_builder=rt.BVH_GPU)
    warm = rt.Scene(path, bvh_builder=rt.BVH_GPU)
    host = rt.Scene(path, bvh_builder=rt.BVH_HOST)
    gi, hi = warm.info(), host.info()
    gn, go = warm.bvh()
    hn, ho = host.bvh()
    refit = refit_bench(warm, host)
    return {
        "triangles": gi.num_triangles,
        "nodes": gi.num_bvh_nodes,
        "host_ms": round(hi.bvh_build_ms, 1),
        "gpu_ms": round(gi.bvh_build_ms, 1),
        "gpu_phase_ms": {k: round(v, 1) for k, v in zip(("setup_upload", "levels", "numbering_download", "host"),
                                                        gi.bvh_gpu_phase_ms)},
        "gpu_mtris_s": round(gi.num_triangles / max(1e-9, gi.bvh_build_ms) / 1e3, 1),
        "identical": bool(np.array_equal(gn, hn) and np.array_equal(go, ho)),
        "refit": refit,
    }


def refit_bench(gpu_scene, host_scene, reps=5):
    """Refit of the same tree to moved vertices (every height scaled by 1.1,
    alternating with the loaded positions): wall time of rt_scene_refit with
    the scene resident on the GPU (median of `reps`) and on the host, and
    whether the two refits give the same nodes."""
    gpu_scene.trace(np.array([[0, 0, 0, 0, 0, 0, 1, 1e8]], np.float32))  # make it resident
    pos, _, _, _ = gpu_scene.triangles()
    moved = pos.copy()
    moved[:, 1::3] *= np.float32(1.1)
    gpu_ms, phases = [], []
    for k in range(reps):
        gpu_scene.refit(moved if k % 2 == 0 else pos)
        i = gpu_scene.info()
        gpu_ms.append(i.refit_ms)
        phases.append(list(i.refit_phase_ms))
    gpu_scene.refit(moved)
    host_scene.refit(moved)
    hi = host_scene.info()
    k = int(np.argsort(gpu_ms)[len(gpu_ms) // 2])
    return {
        "gpu_ms": round(gpu_ms[k], 2),
        "gpu_phase_ms": {n: round(v, 2) for n, v in zip(("host_arrays", "uploads", "node_boxes"), phases[k])},
        "host_ms": round(hi.refit_ms, 2),
        "host_phase_ms": {n: round(v, 2) for n, v in zip(("host_arrays", "uploads", "node_boxes"),
                                                          hi.refit_phase_ms)},
        "identical": bool(np.array_equal(gpu_scene.bvh()[0], host_scene.bvh()[0])),
    }


def textured_bench():
    """Cost of diffuse textures: the textured Cornell scene (PNG albedo on diffuse,
    GGX, subsurface and dielectric materials) at 512x512, 64 spp, depth 12, against
    the same scene with every diffuse_texture key removed."""
    import cudaraytracer_amd as rt
    from cudaraytracer_amd import scenes

    tmp = tempfile.mkdtemp(prefix="rttex_")
    spec = scenes.textured_cornell(width=512, height=512, spp=64, depth=12)
    with_tex = scenes.write_scene(spec, os.path.join(tmp, "tex"), "textured")
    for m in spec.materials:
        m.pop("diffuse_texture", None)
    without = scenes.write_scene(spec, os.path.join(tmp, "plain"), "untextured")
    out = {"workload": "textured_cornell 2596 tris, 5 PNG textures, 512x512, 64 spp, depth 12"}
    for name, path in (("textured", with_tex), ("untextured", without)):
        r = rt.Renderer(rt.Scene(path))
        r.render(64)
        r.synchronize()
        best = None
        for _ in range(3):
            r.reset()
            r.reset_stats()
            t0 = time.perf_counter()
            r.render(64)
            r.synchronize()
            dt = time.perf_counter() - t0
            st = r.stats()
            mr = (st.extension_rays + st.shadow_rays) / dt / 1e6
            best = mr if best is None else max(best, mr)
        out[name + "_mrays_s"] = round(best, 1)
    out["textured_over_untextured"] = round(out["textured_mrays_s"] / out["untextured_mrays_s"], 3)
    return out


def bvh_only():
    tmp = tempfile.mkdtemp(prefix="rtbvh_")
    path = _generate([("terrain_env", {}, tmp, "terrain_bvh")])[0]
    return bvh_build_bench(path)


def _write_scene_job(job):
    """Worker: build one generated scene and write it (JSON + OBJ + PFM)."""
    kind, kwargs, out_dir, name = job
    from cudaraytracer_amd import scenes

    spec = getattr(scenes, kind)(**kwargs)
    return scenes.write_scene(spec, out_dir, name)


def _generate(jobs):
    """Write generated scenes in parallel child interpreters (no GPU state), each a
    `bench.py --write-scenes` run over every workers-th job.  Plain child processes
    rather than a multiprocessing pool: a spawn-context pool leaves its resource
    tracker process running after the benchmark ends."""
    import subprocess

    workers = max(1, min(16, os.cpu_count() or 1, len(jobs)))
    shares = [jobs[k::workers] for k in range(workers)]
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--write-scenes", json.dumps(sh)],
                              stdout=subprocess.PIPE) for sh in shares]
    paths = [None] * len(jobs)
    for k, p in enumerate(procs):
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"scene writer {k} exited with status {p.returncode}")
        paths[k::workers] = json.loads(out)
    return paths


def _load(paths):
    """Load scenes on host threads (the C++ loader and BVH build run without the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    import cudaraytracer_amd as rt

    rt.lib()
    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        return list(ex.map(rt.Scene, paths))


def batched_bench(n: int = 64, seed: int = 4, queue_gib: int = 16):
    """The uneven batch: n seeded scenes (triangles log-uniform 1k-500k, square
    resolution from {128..1024}, spp 1..64, depth 12), all submitted at t = 0.
    Batched: the pool's continuous batching, every stage launch shared by all
    jobs with work left.  Plain: one job at a time, a fresh single-scene
    renderer each, all of a job's samples in one pass (and again with the ray
    sort on, and with one sample per pass as the reference's render loop
    takes them).  Images are bit-identical between all of them (same RNG keys)."""
    import cudaraytracer_amd as rt

    rng = np.random.default_rng(seed)
    jobs, meta = [], []
    tmp = tempfile.mkdtemp(prefix="rtbatch_")
    for k in range(n):
        tris = int(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        width = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.integers(1, 65))
        name = f"batch{k:02d}"
        jobs.append(("batch_scene", {"seed": 1000 + k, "target_tris": tris, "width": width, "spp": spp},
                     os.path.join(tmp, name), name))
        meta.append((width, spp))
    t0 = time.perf_counter()
    paths = _generate(jobs)
    t1 = time.perf_counter()
    sc = _load(paths)
    t2 = time.perf_counter()
    pool = rt.Pool(queue_budget_bytes=queue_gib << 30)
    slots = [pool.add_scene(s) for s in sc]
    reqs = [(0.0, rt.Job(slot, spp=spp)) for slot, (_, spp) in zip(slots, meta)]
    tris_total = sum(s.info().num_triangles for s in sc)
    paths_total = sum(w * w * spp for w, spp in meta)
    # Autotune the pool on a sample of the mix (every 4th job at its full spp:
    # passes as full as the real ones -- a low-spp sample makes small passes,
    # where half-size grids win, while the full mix loses a quarter of its
    # throughput with them), which also warms up kernels and graphs; then one
    # job through the plain loop.
    sample = [rt.Job(slot, spp=spp) for slot, (_, spp) in list(zip(slots, meta))[::4]]
    tune = pool.autotune(sample)
    print(f"[bench] batched autotune: {tune}", file=sys.stderr, flush=True)
    pool.serve([reqs[0]], batched=False)
    pool.reset_stats()
    rep_b, _, job_b = pool.serve(reqs, batched=True)
    st = pool.stats()
    rays = st.extension_rays + st.shadow_rays
    rep_p, _, job_p = pool.serve(reqs, batched=False)
    # The plain loop again with the ray sort on in every single-scene renderer:
    # the ratio below is against the faster of the two plain runs.
    reqs_s = [(0.0, rt.Job(slot, spp=spp, render={"ray_sort": 1})) for slot, (_, spp) in zip(slots, meta)]
    rep_ps, _, _ = pool.serve(reqs_s, batched=False)
    plain_best = max(rep_p["jobs_per_s"], rep_ps["jobs_per_s"])
    # And in the reference's RenderLoop structure: one sample of every pixel
    # per pass (one frame per sample), each job still alone on the GPU.
    reqs_1 = [(0.0, rt.Job(slot, spp=spp, render={"batch_spp": 1})) for slot, (_, spp) in zip(slots, meta)]
    rep_p1, _, _ = pool.serve(reqs_1, batched=False)
    return {
        "workload": f"{n} seeded scenes: Cornell shell + GGX/dielectric/diffuse icospheres + heightfield patch, "
                    "triangles log-uniform 1k-500k, resolution from {128,256,512,1024}^2, spp 1..64, depth 12",
        "triangles_total": int(tris_total),
        "paths_total": int(paths_total),
        "scene_write_s": round(t1 - t0, 2),
        "scene_load_s": round(t2 - t1, 2),
        "batched": {"jobs_per_s": round(rep_b["jobs_per_s"], 3), "makespan_s": round(rep_b["makespan_s"], 4),
                    "job_p50_ms": round(rep_b["job_p50_ms"], 2), "job_p99_ms": round(rep_b["job_p99_ms"], 2),
                    "mrays_s": round(rays / rep_b["makespan_s"] / 1e6, 1)},
        "plain": {"jobs_per_s": round(rep_p["jobs_per_s"], 3), "makespan_s": round(rep_p["makespan_s"], 4),
                  "job_p50_ms": round(rep_p["job_p50_ms"], 2), "job_p99_ms": round(rep_p["job_p99_ms"], 2),
                  "mrays_s": round(rays / rep_p["makespan_s"] / 1e6, 1)},
        "plain_ray_sort": {"jobs_per_s": round(rep_ps["jobs_per_s"], 3), "makespan_s": round(rep_ps["makespan_s"], 4)},
        "plain_one_sample_per_pass": {"jobs_per_s": round(rep_p1["jobs_per_s"], 3),
                                      "makespan_s": round(rep_p1["makespan_s"], 4),
                                      "mrays_s": round(rays / rep_p1["makespan_s"] / 1e6, 1)},
        "speedup": round(rep_b["jobs_per_s"] / plain_best, 3),
        "speedup_vs_one_sample_per_pass": round(rep_b["jobs_per_s"] / rep_p1["jobs_per_s"], 3),
        "queue_gib": queue_gib,
        "autotune": tune,
    }


def serving_bench(rates=(5.0, 10.0, 20.0, 50.0, 100.0, 200.0), n_jobs: int = 30, frames: int = 10, width: int = 1280,
                  height: int = 720, seed: int = 5, queue_gib: int = 16, caps=(4, 8, 16), cap_rates=(100.0, 200.0)):
    """Poisson stream of real-time jobs on terrain + SSS + env scenes (2 M,
    1 M and 0.5 M triangles): each job a `frames`-frame orbit at width x height,
    1 spp per frame, SVGF with 5 a-trous passes.  For each mean rate the same
    arrival sequence is served with continuous batching, with the plain loop
    (one job at a time on a fresh single-scene renderer) and with the plain
    loop on resident renderers (one kept per scene, its history reset per
    job); at `cap_rates` the batched loop also runs with at most `caps`
    real-time jobs admitted at once.  p50/p99 of frame and job latency and
    jobs/s are reported."""
    import cudaraytracer_amd as rt

    tmp = tempfile.mkdtemp(prefix="rtserve_")
    specs = [("terrain_env", {"seed": sd, "grid": g, "width": width, "height": height, "spp": 1},
              os.path.join(tmp, f"terrain{sd}"), f"terrain{sd}") for sd, g in ((3, 1001), (5, 709), (7, 501))]
    sc = _load(_generate(specs))
    # Continuous batching with cohort ticks: one cohort's frames traced while the
    # other's are filtered on two side streams (rt_pool_params.cohort_ticks).
    pool = rt.Pool(queue_budget_bytes=queue_gib << 30, pipelined=True)
    slots = [pool.add_scene(s) for s in sc]
    base_cams = [rt.Renderer(s, batch_spp=1).camera() for s in sc]
    out = {"workload": f"Poisson arrivals of {n_jobs} jobs per rate; each a {frames}-frame orbit at "
                       f"{width}x{height}, 1 spp/frame, SVGF (5 a-trous); scenes: terrain_env 2.0M/1.0M/0.5M "
                       "triangles + SSS sphere + 2048x1024 env sampled by alias tables", "rates": []}
    # Tune on passes shaped like the saturated ones (24 concurrent jobs): a
    # 6-job sample makes small passes, where half-size grids win, and its pick
    # cost the saturated stream ~8 % (measured).
    sample = []
    for k in range(24):
        c = base_cams[k % len(sc)]
        cams = [rt.Camera(tuple(c.target), c.distance, c.theta_deg, c.phi_deg + 0.5 * f) for f in range(2)]
        sample.append(rt.Job(slots[k % len(sc)], spp=1, frames=2, cameras=cams, render={"env_sampling": 1}))
    out["autotune"] = pool.autotune(sample)
    print(f"[bench] serving autotune: {out['autotune']}", file=sys.stderr, flush=True)
    tuning = pool.tuning()
    warm = rt.Job(slots[2], spp=1, frames=2, cameras=[base_cams[2]] * 2)
    pool.serve([(0.0, warm)], batched=False)
    pool.serve([(0.0, warm)], batched="resident")

    def stream(rate):
        rng = np.random.default_rng(seed + int(rate))
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
        arrivals -= arrivals[0]
        reqs = []
        for k, t in enumerate(arrivals):
            s = int(rng.integers(0, len(sc)))
            c = base_cams[s]
            phi0 = float(rng.uniform(-20, 20))
            cams = [rt.Camera(tuple(c.target), c.distance, c.theta_deg, c.phi_deg + phi0 + 0.5 * f)
                    for f in range(frames)]
            reqs.append((float(t), rt.Job(slots[s], spp=1, frames=frames, cameras=cams, render={"env_sampling": 1})))
        return reqs

    rnd = lambda rep: {k: round(v, 3) if isinstance(v, float) else v for k, v in rep.items()}
    for rate in rates:
        reqs = stream(rate)
        row = {"rate_jobs_s": rate}
        for mode, arg in (("batched", True), ("plain", False), ("plain_resident", "resident")):
            rep, _, _ = pool.serve(reqs, batched=arg)
            row[mode] = rnd(rep)
            print(f"[bench] serving rate {rate}: {mode} {row[mode]}", file=sys.stderr, flush=True)
        row["jobs_per_s_ratio"] = round(row["batched"]["jobs_per_s"] / row["plain"]["jobs_per_s"], 3)
        row["jobs_per_s_ratio_vs_resident"] = round(row["batched"]["jobs_per_s"] / row["plain_resident"]["jobs_per_s"],
                                                    3)
        out["rates"].append(row)
    # Capped admission: the same streams with at most `cap` real-time jobs at once.
    out["capped"] = []
    for cap in caps:
        cp = rt.Pool(queue_budget_bytes=queue_gib << 30, max_realtime_jobs=cap, pipelined=True)
        for s in sc:
            cp.add_scene(s)
        cp.set_tuning(tuning.grid_scale, tuning.pass_paths, tuning.tile_width, bool(tuning.ray_sort))
        cp.serve([(0.0, warm)], batched=True)
        for rate in cap_rates:
            rep, _, _ = cp.serve(stream(rate), batched=True)
            row = {"max_realtime_jobs": cap, "rate_jobs_s": rate, "batched": rnd(rep)}
            print(f"[bench] serving cap {cap} rate {rate}: {row['batched']}", file=sys.stderr, flush=True)
            out["capped"].append(row)
        cp.close()
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene_path: str, spp: int, full: bool = False, pixel_stride: int = 1, scene=None,
                 env_alias: bool = False, runs: int = 3):
    """The CPU oracle (the reference integrator restated in C, OpenMP over
    pixels, every host core) on the same scene: the whole image (`full`) or
    every `pixel_stride`-th pixel, `spp` samples each; median of `runs`.  With
    `scene` (a loaded library scene whose objects import as one mesh each) the
    oracle takes the imported triangles instead of parsing the OBJ files
    itself (its Python reader is slow at millions of triangles)."""
    from oracle import rtref

    rtref.build()
    geometry = None
    if scene is not None:
        pos, nrm, uv, _ = scene.triangles()
        mesh4, _ = scene.meshes()
        with open(scene_path, encoding="utf-8-sig") as f:
            objects = json.load(f)["objects"]
        geometry = rtref.geometry_from_arrays(objects, pos.reshape(-1, 9), nrm.reshape(-1, 9), uv.reshape(-1, 6),
                                              mesh4[:, 2], mesh4[:, 3])
    t_setup = time.perf_counter()
    a = rtref.parse_scene(scene_path, geometry)
    rs = rtref.RefScene(a)
    setup_s = time.perf_counter() - t_setup
    p = rs.params(env_alias=env_alias)
    npx = a.width * a.height
    pixels = None if full or pixel_stride <= 1 else np.arange(0, npx, pixel_stride, dtype=np.int32)
    rs.render(1, p, pixels=np.arange(0, npx, 997, dtype=np.int32))  # warm-up
    rates, secs = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        _, _, st = rs.render(spp, p, pixels=pixels)
        dt = time.perf_counter() - t0
        rays = int(st[0] + st[1])
        rates.append(rays / dt / 1e6)
        secs.append(dt)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = f"{a.width}x{a.height}" if pixels is None else f"every {pixel_stride}th pixel of {a.width}x{a.height}"
    return {"value": round(float(np.median(rates)), 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"median of {runs} runs of {what} x {spp} spp of the same scene, environment sampling "
                      f"{'alias tables' if env_alias else 'CDF search'} ({sum(secs):.1f} s total, {rays} rays per "
                      f"run; oracle built -O3 -fopenmp -ffp-contract=off; scene set-up {setup_s:.1f} s untimed)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-pixel-stride", type=int, default=4,
                    help="the CPU baseline renders every Nth pixel of the headline workload at its full spp")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-file", default=C3_TRAFFIC,
                    help="HBM bytes per launch measured with rocprofv3 --pmc (table from --traffic-from-pmc)")
    ap.add_argument("--only", choices=("frames", "c1", "c2", "batched", "serving", "bvh", "textured"), default=None,
                    help="run one secondary workload and print its JSON")
    ap.add_argument("--calibrate-copy", action="store_true", help="run the 1 GiB copy that scales the counters")
    ap.add_argument("--write-scenes", metavar="JOBS_JSON", help=argparse.SUPPRESS)
    ap.add_argument("--kernel-table", metavar="DB", help="print a rocprofv3 --kernel-trace database as a table")
    ap.add_argument("--counter-table", nargs=2, metavar=("sq|l2", "DB"),
                    help="print a rocprofv3 --pmc database (SQ or L2/lane counters) as a table")
    ap.add_argument("--traffic-from-pmc", nargs=4, metavar=("FETCH_DB", "WRITE_DB", "COPY_FETCH_DB", "COPY_WRITE_DB"),
                    help="print the traffic table from rocprofv3 --pmc databases")
    ap.add_argument("--no-frames", action="store_true", help="skip the real-time SVGF frame-loop measurement")
    ap.add_argument("--no-secondary", action="store_true", help="skip the C1 / C2 / BVH-build sections")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-scene batched-vs-plain measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the Poisson serving measurement")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU is design-only)"}))
        return 2
    if args.write_scenes:
        print(json.dumps([_write_scene_job(tuple(job)) for job in json.loads(args.write_scenes)]))
        return 0
    if args.kernel_table:
        print(kernel_table(args.kernel_table))
        return 0
    if args.counter_table:
        print(counter_table(args.counter_table[0], args.counter_table[1]))
        return 0
    if args.traffic_from_pmc:
        print(traffic_from_pmc(*args.traffic_from_pmc))
        return 0
    if args.calibrate_copy:
        calibrate_copy()
        return 0
    if args.only:
        fn = {"frames": frames_bench, "c1": c1_bench, "c2": lambda: c2_bench()[0], "batched": batched_bench,
              "serving": serving_bench, "bvh": bvh_only, "textured": textured_bench}[args.only]
        print(json.dumps({args.only: fn()}))
        return 0

    import cudaraytracer_amd as rt

    path = c3_scene_path()
    scene = rt.Scene(path)
    info = scene.info()
    print(f"[bench] scene loaded: {info.num_triangles} tris, {info.width}x{info.height}, {info.spp} spp, "
          f"depth {info.max_bounces}", file=sys.stderr, flush=True)
    m = measure_render(scene, args.steps, args.warmup, args.traffic_file, env_alias=True)
    print(f"[bench] {args.steps} steps: {m['ms_per_step']:.2f} ms/step, {m['mrays_s']:.1f} Mrays/s",
          file=sys.stderr, flush=True)
    dom = m.pop("dominant")
    out = {
        "metric": "Mrays/s & % HBM roofline; served jobs/s + p50/p99 latency vs plain",
        "value": m["mrays_s"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene generated in the reference's scene-JSON + OBJ format)",
        "config": {"workload": "terrain_env: 2.0M-triangle fBm terrain + random-walk subsurface icosphere (5120 "
                               "tris, sigma_a and sigma_s log-uniform 0.1-10 per m) under a procedural 2048x1024 "
                               "FP32 HDR sky (Gaussian sun lobe, peak 1e4) sampled by alias tables, 1920x1080, "
                               "32 spp, depth 16",
                   "triangles": info.num_triangles, "width": info.width, "height": info.height, "spp": info.spp,
                   "max_depth": info.max_bounces, "batch_spp": m["batch_spp"], "rays_per_step": m["rays_per_step"],
                   "ray_sort": False},
        "roofline": {k: dom[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_kernel": dom["kernel"],
        "stages": m["stages"],
        "stage_ms_per_step": m["stage_ms_per_step"],
        "nodes_per_ext_ray": m["nodes_per_ext_ray"],
        "tris_per_ext_ray": m["tris_per_ext_ray"],
        "path_counts_per_step": m["path_counts"],
    }
    # The shadow test's other modes on the same workload (rt_render_params.shadow_any_hit):
    # 1 = early exit, whose full render must equal mode 0's bit for bit; 2 = the segment test,
    # an approximation, whose full render is compared with mode 0's.
    acc = {}
    for mode in (0, 1, 2):
        rr = rt.Renderer(scene, env_alias=True, shadow_any_hit=mode)
        rr.render(info.spp)
        acc[mode] = [rr.read_accum(k) for k in (rt.RT_ACCUM_DIRECT, rt.RT_ACCUM_INDIRECT, rt.RT_ACCUM_MEAN)]
        del rr
    out["shadow_modes"] = {}
    for mode, name in ((1, "early_exit"), (2, "segment")):
        mm = measure_render(scene, args.steps, args.warmup, args.traffic_file, env_alias=True, shadow_any_hit=mode)
        mm.pop("dominant")
        shadow = dict(mm["stages"]["k_shadow"], traffic=None)  # the measured traffic is mode 0's
        rec = {"shadow_any_hit": mode, "mrays_s": mm["mrays_s"], "ms_per_step": mm["ms_per_step"],
               "stage_ms_per_step": mm["stage_ms_per_step"], "shadow": shadow}
        if mode == 1:
            rec["image_bit_identical_to_mode_0"] = all(
                np.array_equal(x.view(np.int32), y.view(np.int32)) for x, y in zip(acc[0][:2], acc[1][:2]))
        else:
            ref, img = acc[0][2][..., :3].astype(np.float64), acc[2][2][..., :3].astype(np.float64)
            rec["image_vs_mode_0"] = {
                "mean_rel_diff": round(float((img.mean() - ref.mean()) / ref.mean()), 6),
                "rmse_over_mean": round(float(np.sqrt(np.mean((img - ref) ** 2)) / ref.mean()), 6),
                "pixels_differing": round(float(np.mean(np.any(img != ref, axis=-1))), 6)}
        out["shadow_modes"][name] = rec
        print(f"[bench] shadow mode {mode}: {mm['mrays_s']} Mrays/s, shadow {mm['stage_ms_per_step']['shadow']} ms "
              f"{ {k: v for k, v in rec.items() if k.startswith('image')} }", file=sys.stderr, flush=True)
    del acc
    # The same workload with the reference's subsurface method (Burley profile and
    # two probe rays, parameters mapped from the medium's coefficients) instead
    # of the random walk: the scene JSON rewritten beside the original.
    bpath = burley_variant(path)
    mb = measure_render(rt.Scene(bpath), args.steps, args.warmup, None, env_alias=True)
    mb.pop("dominant")
    out["sss_burley"] = {"subsurface_method": "burley", "mrays_s": mb["mrays_s"], "ms_per_step": mb["ms_per_step"],
                         "stage_ms_per_step": mb["stage_ms_per_step"],
                         "subsurface_paths_per_step": mb["path_counts"]["shaded_hits"]["subsurface"]}
    print(f"[bench] Burley subsurface variant: {mb['mrays_s']} Mrays/s, {mb['ms_per_step']} ms/step",
          file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle on a pixel sample of the same workload)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(path, info.spp, pixel_stride=args.cpu_pixel_stride, scene=scene,
                                           env_alias=True)
        print(f"[bench] cpu_baseline: {out['cpu_baseline']}", file=sys.stderr, flush=True)
    del scene
    if not args.no_secondary:
        print("[bench] C2: spheres 512x512 64 spp", file=sys.stderr, flush=True)
        out["c2"], c2_path = c2_bench()
        print("[bench] C1: Cornell box 256x256 16 spp, GPU and CPU", file=sys.stderr, flush=True)
        out["c1"] = c1_bench(cpu=not args.no_cpu_baseline)
        print("[bench] BVH build and refit of the terrain", file=sys.stderr, flush=True)
        out["c3_extra"] = c3_extra(path)
    if not args.no_frames:
        print("[bench] real-time frame loop (SVGF)", file=sys.stderr, flush=True)
        out["frames"] = frames_bench()
    if not args.no_batched:
        print("[bench] batched: 64 uneven scenes, pool vs one at a time", file=sys.stderr, flush=True)
        out["batched"] = batched_bench()
        print(f"[bench] batched: {out['batched']}", file=sys.stderr, flush=True)
    if not args.no_serving:
        print("[bench] serving: Poisson stream of real-time jobs", file=sys.stderr, flush=True)
        out["serving"] = serving_bench()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
