"""MI355X wavefront path tracer: Python mirror of the reference's Scene / Renderer.

The reference drives rendering as ``Scene scene(path); Renderer renderer(ui);
renderer.RenderLoop();`` (Src/main.cu:26-33).  Here the same three steps are

    scene = Scene("scene.json")            # Scene::Scene (Src/Core/scene.cpp:50-57)
    r = Renderer(scene)                     # Renderer::Renderer (Src/Core/renderer.cuh:1094-1100)
    r.render(spp)                           # RenderLoop frames, one sample per pixel each
    img = r.read_accum(RT_ACCUM_MEAN)       # renderBuffers read-back

and, for the real-time denoised loop (RenderLoop with denoise on,
Src/Core/renderer.cuh:1141-1262),

    r.set_camera(Camera(...))               # orbit-camera move (Src/Utils/ui.cpp:313-348)
    r.render_frame(spp=1)                   # G-buffer, path trace, SVGF
    img = r.read_frame(RT_FRAME_DENOISED)

Everything runs in the HIP library ``librtamd.so`` through its C ABI
(include/rt.h); there is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtamd.so")

RT_OK = 0
RT_ACCUM_DIRECT = 0
RT_ACCUM_INDIRECT = 1
RT_ACCUM_MEAN = 2
# Gather modes (rt_read_gather).
RT_GATHER_COLOR, RT_GATHER_DEPTH, RT_GATHER_NORMAL, RT_GATHER_ID, RT_GATHER_POSITION, RT_GATHER_DENOISED = range(6)
# Image files (rt_save_image).
RT_IMAGE_PNG_COLOR = 0
RT_IMAGE_PNG_DENOISED = 1
RT_IMAGE_PFM_MEAN = 2
RT_IMAGE_PFM_DENOISED = 3
STAGES = ("raygen", "extend", "shade", "shadow", "resolve", "gbuffer", "temporal", "variance", "atrous", "sort")
# Shading classes (rt_stats_out.class_paths): one shade kernel each; "light" is a
# path that hit an emitter and ends there (misses end in extend and are in no class).
SHADE_CLASSES = ("light", "diffuse", "ggx", "dielectric", "subsurface")

# Frame buffers (rt_read_frame): id -> (numpy dtype, channels).
RT_FRAME_DENOISED = 0
RT_FRAME_DIRECT = 1
RT_FRAME_INDIRECT = 2
RT_FRAME_HISTORY_DIRECT = 3
RT_FRAME_HISTORY_INDIRECT = 4
RT_FRAME_MOMENTS_DIRECT = 5
RT_FRAME_MOMENTS_INDIRECT = 6
RT_FRAME_HISTORY_LENGTH = 7
RT_FRAME_VARIANCE_DIRECT = 8
RT_FRAME_VARIANCE_INDIRECT = 9
RT_FRAME_G_POSITION_TRI = 10
RT_FRAME_G_NORMAL_MAT = 11
RT_FRAME_G_BARY_MESH = 12
RT_FRAME_G_MOTION_DEPTH = 13
FRAME_LAYOUT = {
    RT_FRAME_MOMENTS_DIRECT: (np.float32, 2), RT_FRAME_MOMENTS_INDIRECT: (np.float32, 2),
    RT_FRAME_HISTORY_LENGTH: (np.int32, 1),
    RT_FRAME_G_NORMAL_MAT: (np.uint16, 4), RT_FRAME_G_BARY_MESH: (np.uint16, 4),
}

# C symbols declared in include/rt.h (checked by the test suite).
EXPORTS = (
    "rt_render_params_default", "rt_scene_load", "rt_scene_info_get", "rt_scene_destroy",
    "rt_renderer_create", "rt_renderer_destroy", "rt_render", "rt_render_counted", "rt_synchronize",
    "rt_reset", "rt_reset_stats", "rt_read_accum", "rt_read_rgba8", "rt_stats", "rt_last_error",
    "rt_scene_refit", "rt_scene_get_triangles", "rt_scene_get_meshes", "rt_scene_get_bvh", "rt_scene_get_texture",
    "rt_scene_get_envmap", "rt_scene_get_materials", "rt_scene_get_medium", "rt_trace_rays_noalloc",
    "rt_texture_lookup", "rt_trace_rays", "rt_device_math", "rt_save_image", "rt_read_gather", "rt_read_denoised", "rt_read_gbuffer",
    "rt_denoise_params_default", "rt_get_camera", "rt_set_camera", "rt_render_frame", "rt_reset_history",
    "rt_read_frame", "rt_pool_create", "rt_pool_destroy", "rt_pool_add_scene", "rt_pool_submit", "rt_pool_step",
    "rt_pool_job_status", "rt_pool_read_accum", "rt_pool_read_frame", "rt_pool_release", "rt_pool_stats",
    "rt_pool_reset_stats", "rt_pool_serve", "rt_pool_set_tuning", "rt_pool_get_tuning", "rt_pool_autotune",
)
RT_ERR_BUSY = -7


class RtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rt error {code}: {msg}")
        self.code = code


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("num_triangles", ctypes.c_int), ("num_meshes", ctypes.c_int), ("num_materials", ctypes.c_int),
        ("num_lights", ctypes.c_int), ("num_bvh_nodes", ctypes.c_int), ("bvh_max_depth", ctypes.c_int),
        ("bvh_fallback_splits", ctypes.c_int), ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("spp", ctypes.c_int), ("max_bounces", ctypes.c_int), ("rr", ctypes.c_float), ("denoise", ctypes.c_int),
        ("seed", ctypes.c_uint32), ("rr_start_depth", ctypes.c_int),
        ("bounds_lo", ctypes.c_float * 3), ("bounds_hi", ctypes.c_float * 3),
        ("camera_position", ctypes.c_float * 3), ("camera_left_down", ctypes.c_float * 3),
        ("camera_du", ctypes.c_float * 3), ("camera_dv", ctypes.c_float * 3), ("num_textures", ctypes.c_int),
        ("bvh_build_ms", ctypes.c_float), ("bvh_on_gpu", ctypes.c_int), ("bvh_gpu_phase_ms", ctypes.c_float * 4),
        ("load_phase_ms", ctypes.c_float * 3), ("refit_ms", ctypes.c_float), ("refit_on_gpu", ctypes.c_int),
        ("refit_phase_ms", ctypes.c_float * 3),
    ]


class RenderParams(ctypes.Structure):
    _fields_ = [
        ("max_bounces", ctypes.c_int), ("rr", ctypes.c_float), ("rr_start_depth", ctypes.c_int),
        ("seed", ctypes.c_int64), ("batch_spp", ctypes.c_int), ("queue_budget_bytes", ctypes.c_int64),
        ("flush_survivors", ctypes.c_int), ("profile", ctypes.c_int), ("env_sampling", ctypes.c_int),
        ("disable_graphs", ctypes.c_int), ("ray_sort", ctypes.c_int), ("shadow_any_hit", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("extension_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("extension_node_visits", ctypes.c_uint64), ("extension_tri_tests", ctypes.c_uint64),
        ("shadow_node_visits", ctypes.c_uint64), ("shadow_tri_tests", ctypes.c_uint64),
        ("probe_rays", ctypes.c_uint64), ("stage_ms", ctypes.c_double * len(STAGES)),
        ("stage_launches", ctypes.c_int * len(STAGES)),
        ("spp_done", ctypes.c_int), ("batch_spp", ctypes.c_int),
        ("miss_paths", ctypes.c_uint64), ("shadow_lit", ctypes.c_uint64),
        ("class_paths", ctypes.c_uint64 * len(SHADE_CLASSES)),
    ]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_
             if f not in ("stage_ms", "stage_launches", "class_paths")}
        d["stage_ms"] = {s: self.stage_ms[i] for i, s in enumerate(STAGES)}
        d["class_paths"] = {c: self.class_paths[i] for i, c in enumerate(SHADE_CLASSES)}
        d["stage_launches"] = {s: self.stage_launches[i] for i, s in enumerate(STAGES)}
        return d


class Camera(ctypes.Structure):
    """Orbit camera: target, distance, polar angle theta from +y, azimuth phi (degrees)."""
    _fields_ = [("target", ctypes.c_float * 3), ("distance", ctypes.c_float),
                ("theta_deg", ctypes.c_float), ("phi_deg", ctypes.c_float)]

    def __repr__(self) -> str:
        return (f"Camera(target={list(self.target)}, distance={self.distance}, "
                f"theta_deg={self.theta_deg}, phi_deg={self.phi_deg})")


class DenoiseParams(ctypes.Structure):
    """DenoiseParam (Src/Utils/ui.h:14-25); defaults from rt_denoise_params_default."""
    _fields_ = [("max_history", ctypes.c_int), ("sigma_light", ctypes.c_float), ("sigma_normal", ctypes.c_float),
                ("sigma_depth", ctypes.c_float), ("reprojection_depth", ctypes.c_float),
                ("atrous_iterations", ctypes.c_int)]

    @classmethod
    def default(cls, **kw) -> "DenoiseParams":
        p = cls()
        lib().rt_denoise_params_default(ctypes.byref(p))
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class PoolParams(ctypes.Structure):
    _fields_ = [("queue_budget_bytes", ctypes.c_int64), ("hbm_budget_bytes", ctypes.c_int64), ("profile", ctypes.c_int),
                ("cohort_ticks", ctypes.c_int), ("max_realtime_jobs", ctypes.c_int)]


class JobDescC(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_int), ("spp", ctypes.c_int), ("frames", ctypes.c_int),
                ("cameras", ctypes.POINTER(Camera)), ("render", ctypes.POINTER(RenderParams)),
                ("denoise", ctypes.POINTER(DenoiseParams))]


class JobStatus(ctypes.Structure):
    _fields_ = [("done", ctypes.c_int), ("spp_done", ctypes.c_int), ("frames_done", ctypes.c_int),
                ("submit_s", ctypes.c_double), ("start_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class PoolTuning(ctypes.Structure):
    _fields_ = [("grid_scale", ctypes.c_float), ("pass_paths", ctypes.c_int64), ("tile_width", ctypes.c_int),
                ("ray_sort", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuneReport(ctypes.Structure):
    _fields_ = [("best", PoolTuning), ("baseline_ms", ctypes.c_double), ("best_ms", ctypes.c_double),
                ("runs", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {"best": self.best.as_dict(), "baseline_ms": self.baseline_ms, "best_ms": self.best_ms,
                "runs": self.runs}


class ServeRequestC(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("job", JobDescC), ("result_rgba", ctypes.POINTER(ctypes.c_float))]


class ServeReport(ctypes.Structure):
    _fields_ = [("jobs", ctypes.c_int), ("frames", ctypes.c_int), ("makespan_s", ctypes.c_double),
                ("jobs_per_s", ctypes.c_double), ("frame_p50_ms", ctypes.c_double), ("frame_p99_ms", ctypes.c_double),
                ("job_p50_ms", ctypes.c_double), ("job_p99_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_LIB: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Load librtamd.so (built in-tree by ``python -m cudaraytracer_amd.build``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m cudaraytracer_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, fp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float)
    L.rt_last_error.restype = ctypes.c_char_p
    L.rt_render_params_default.argtypes = [ctypes.POINTER(RenderParams)]
    L.rt_render_params_default.restype = None
    L.rt_scene_load.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(vp)]
    L.rt_scene_info_get.argtypes = [vp, ctypes.POINTER(SceneInfo)]
    L.rt_scene_destroy.argtypes = [vp]
    L.rt_scene_destroy.restype = None
    L.rt_renderer_create.argtypes = [vp, ctypes.POINTER(RenderParams), ctypes.POINTER(vp)]
    L.rt_renderer_destroy.argtypes = [vp]
    L.rt_renderer_destroy.restype = None
    for name in ("rt_render", "rt_render_counted"):
        getattr(L, name).argtypes = [vp, ip]
    for name in ("rt_synchronize", "rt_reset", "rt_reset_stats"):
        getattr(L, name).argtypes = [vp]
    L.rt_read_accum.argtypes = [vp, ip, fp]
    L.rt_read_rgba8.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8)]
    L.rt_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.rt_scene_refit.argtypes = [vp, fp, fp, ctypes.c_int64]
    L.rt_scene_get_triangles.argtypes = [vp, fp, fp, fp, ctypes.POINTER(ctypes.c_int32)]
    L.rt_scene_get_meshes.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), fp]
    L.rt_scene_get_bvh.argtypes = [vp, ip, vp, ctypes.POINTER(ctypes.c_int32)]
    L.rt_scene_get_texture.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_uint8)]
    L.rt_texture_lookup.argtypes = [vp, ip, fp, ip, fp]
    L.rt_scene_get_materials.argtypes = [vp, ctypes.POINTER(ctypes.c_int), vp]
    L.rt_scene_get_medium.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), vp, vp, vp]
    L.rt_scene_get_envmap.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), fp]
    L.rt_save_image.argtypes = [vp, ip, ctypes.c_char_p]
    L.rt_read_gather.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_uint8)]
    L.rt_read_denoised.argtypes = [vp, fp]
    L.rt_read_gbuffer.argtypes = [vp, ip, vp]
    L.rt_trace_rays.argtypes = [vp, fp, ip, fp, ctypes.POINTER(ctypes.c_uint32)]
    L.rt_trace_rays_noalloc.argtypes = [vp, fp, ip, fp, ctypes.POINTER(ctypes.c_uint32)]
    L.rt_device_math.argtypes = [ip, fp, fp, fp, ip]
    L.rt_denoise_params_default.argtypes = [ctypes.POINTER(DenoiseParams)]
    L.rt_denoise_params_default.restype = None
    L.rt_get_camera.argtypes = [vp, ctypes.POINTER(Camera)]
    L.rt_set_camera.argtypes = [vp, ctypes.POINTER(Camera)]
    L.rt_render_frame.argtypes = [vp, ip, ctypes.POINTER(DenoiseParams)]
    L.rt_reset_history.argtypes = [vp]
    L.rt_read_frame.argtypes = [vp, ip, vp, ctypes.c_size_t]
    L.rt_pool_create.argtypes = [ctypes.POINTER(PoolParams), ctypes.POINTER(vp)]
    L.rt_pool_destroy.argtypes = [vp]
    L.rt_pool_destroy.restype = None
    L.rt_pool_add_scene.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int)]
    L.rt_pool_submit.argtypes = [vp, ctypes.POINTER(JobDescC), ctypes.POINTER(ctypes.c_int)]
    L.rt_pool_step.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.rt_pool_job_status.argtypes = [vp, ip, ctypes.POINTER(JobStatus)]
    L.rt_pool_read_accum.argtypes = [vp, ip, ip, fp]
    L.rt_pool_read_frame.argtypes = [vp, ip, ip, vp, ctypes.c_size_t]
    L.rt_pool_release.argtypes = [vp, ip]
    L.rt_pool_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.rt_pool_reset_stats.argtypes = [vp]
    L.rt_pool_set_tuning.argtypes = [vp, ctypes.POINTER(PoolTuning)]
    L.rt_pool_get_tuning.argtypes = [vp, ctypes.POINTER(PoolTuning)]
    L.rt_pool_autotune.argtypes = [vp, ctypes.POINTER(JobDescC), ip, ctypes.POINTER(TuneReport)]
    L.rt_pool_serve.argtypes = [vp, ctypes.POINTER(ServeRequestC), ip, ip, ctypes.POINTER(ServeReport),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    _LIB = L
    return L


def _check(code: int) -> None:
    if code != RT_OK:
        raise RtError(code, lib().rt_last_error().decode(errors="replace"))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class LoadOpts(ctypes.Structure):
    _fields_ = [("bvh_builder", ctypes.c_int)]


BVH_AUTO, BVH_HOST, BVH_GPU = 0, 1, 2


class Scene:
    """A loaded scene: JSON + OBJ (+ environment map) and its SAH BVH, built on
    the host or the GPU (`bvh_builder`: BVH_AUTO, BVH_HOST, BVH_GPU; same tree)."""

    def __init__(self, json_path: str, bvh_builder: int = BVH_AUTO):
        self._h = ctypes.c_void_p()
        opts = LoadOpts(bvh_builder)
        _check(lib().rt_scene_load(os.fsencode(json_path), ctypes.byref(opts), ctypes.byref(self._h)))
        self.path = json_path

    def close(self) -> None:
        if self._h:
            lib().rt_scene_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self) -> SceneInfo:
        out = SceneInfo()
        _check(lib().rt_scene_info_get(self._h, ctypes.byref(out)))
        return out

    def triangles(self):
        n = self.info().num_triangles
        pos = np.zeros((n, 9), np.float32)
        nrm = np.zeros((n, 9), np.float32)
        uv = np.zeros((n, 6), np.float32)
        mesh = np.zeros(n, np.int32)
        _check(lib().rt_scene_get_triangles(self._h, _fptr(pos), _fptr(nrm), _fptr(uv), _iptr(mesh)))
        return pos, nrm, uv, mesh

    def refit(self, pos: np.ndarray, nrm: Optional[np.ndarray] = None) -> None:
        """Move the vertices (pos: (n, 9) float32 in scene order, as triangles()
        returns them; nrm likewise or None to keep the shading normals) and
        refit the BVH in place, on the GPU when the scene is resident there."""
        pos = np.ascontiguousarray(pos, np.float32)
        n = pos.size // 9
        if pos.size != 9 * n:
            raise ValueError("pos must hold 9 floats per triangle")
        npt = None
        if nrm is not None:
            nrm = np.ascontiguousarray(nrm, np.float32)
            if nrm.size != pos.size:
                raise ValueError("nrm must match pos")
            npt = _fptr(nrm)
        _check(lib().rt_scene_refit(self._h, _fptr(pos), npt, n))

    def meshes(self):
        m = self.info().num_meshes
        mesh4 = np.zeros((m, 4), np.int32)
        area = np.zeros(m, np.float32)
        _check(lib().rt_scene_get_meshes(self._h, _iptr(mesh4), _fptr(area)))
        return mesh4, area

    def bvh(self, from_packed: bool = False):
        info = self.info()
        nodes = np.zeros((info.num_bvh_nodes, 8), np.int32)
        order = np.zeros(info.num_triangles, np.int32)
        _check(lib().rt_scene_get_bvh(self._h, int(from_packed), nodes.ctypes.data_as(ctypes.c_void_p), _iptr(order)))
        return nodes, order

    def texture(self, index: int) -> np.ndarray:
        """Decoded diffuse texture `index`: (h, w, 4) uint8, row 0 = top."""
        w, h = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().rt_scene_get_texture(self._h, index, ctypes.byref(w), ctypes.byref(h), None))
        out = np.zeros((h.value, w.value, 4), np.uint8)
        _check(lib().rt_scene_get_texture(self._h, index, ctypes.byref(w), ctypes.byref(h),
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def materials(self) -> np.ndarray:
        """Material records in "materials" order: (n, 24) float32 words (type, basecolor,
        emission, roughness, metallic, transmission, ior, scatter distance, texture, valid,
        sigma_a, sigma_s, subsurface method, padding; type, texture, valid and the method
        are int32 bit patterns)."""
        n = ctypes.c_int(0)
        _check(lib().rt_scene_get_materials(self._h, ctypes.byref(n), None))
        out = np.zeros((n.value, 24), np.float32)
        _check(lib().rt_scene_get_materials(self._h, ctypes.byref(n), out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def medium(self, mesh: int):
        """Clearance grid of a random-walk medium: (lo, inv, cells (n, n, n) indexed
        [z, y, x]) or None when the mesh is no such medium."""
        n = ctypes.c_int(0)
        _check(lib().rt_scene_get_medium(self._h, mesh, ctypes.byref(n), None, None, None))
        if n.value == 0:
            return None
        lo, inv = np.zeros(3, np.float32), np.zeros(3, np.float32)
        cells = np.zeros((n.value,) * 3, np.float32)
        _check(lib().rt_scene_get_medium(self._h, mesh, ctypes.byref(n), lo.ctypes.data_as(ctypes.c_void_p),
                                         inv.ctypes.data_as(ctypes.c_void_p), cells.ctypes.data_as(ctypes.c_void_p)))
        return lo, inv, cells

    def envmap(self) -> np.ndarray:
        """Decoded environment map: (h, w, 3) float32, row 0 = the file's first scanline
        (what stbi_loadf returns); (0, 0, 3) without an environment light."""
        w, h = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().rt_scene_get_envmap(self._h, ctypes.byref(w), ctypes.byref(h), None))
        out = np.zeros((h.value, w.value, 3), np.float32)
        _check(lib().rt_scene_get_envmap(self._h, ctypes.byref(w), ctypes.byref(h), _fptr(out)))
        return out

    def texture_lookup(self, index: int, uv: np.ndarray) -> np.ndarray:
        """Albedo of texture `index` at (n, 2) float32 (u, v) pairs, on the GPU: (n, 3) float32."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 3), np.float32)
        _check(lib().rt_texture_lookup(self._h, index, _fptr(uv), uv.shape[0], _fptr(out)))
        return out

    def trace(self, rays: np.ndarray):
        """Closest hits on the GPU. rays: (n, 8) float32 = o.xyz, tmin, d.xyz, tmax."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        counts = np.zeros((n, 2), np.uint32)
        _check(lib().rt_trace_rays(self._h, _fptr(rays), n, _fptr(hits),
                                   counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        prim = hits[:, 3].copy().view(np.int32)
        return hits[:, 0].copy(), hits[:, 1].copy(), hits[:, 2].copy(), prim, counts


class Renderer:
    """Wavefront renderer for one scene; owns its HIP stream and device buffers."""

    def __init__(self, scene: Scene, max_bounces: int = -1, rr: float = -1.0, rr_start_depth: int = -1,
                 seed: int = -1, batch_spp: int = 0, queue_budget_bytes: int = 0, flush_survivors: bool = False,
                 profile: bool = False, env_alias: bool = False, graphs: bool = True, ray_sort: bool = False,
                 shadow_any_hit: int = 0):
        """shadow_any_hit: 0 (or False) the reference's shadow test, 1 (or True) its early-exit
        form (same results), 2 the segment test (fast mode; see rt_render_params)."""
        p = RenderParams()
        lib().rt_render_params_default(ctypes.byref(p))
        p.max_bounces, p.rr, p.rr_start_depth, p.seed = max_bounces, rr, rr_start_depth, seed
        p.batch_spp, p.queue_budget_bytes = batch_spp, queue_budget_bytes
        p.flush_survivors, p.profile = int(flush_survivors), int(profile)
        p.env_sampling = int(env_alias)
        p.disable_graphs = 0 if graphs else 1
        p.ray_sort = int(ray_sort)
        p.shadow_any_hit = int(shadow_any_hit)
        self.scene = scene
        self._h = ctypes.c_void_p()
        _check(lib().rt_renderer_create(scene.handle, ctypes.byref(p), ctypes.byref(self._h)))
        info = scene.info()
        self.width, self.height = info.width, info.height

    def close(self) -> None:
        if self._h:
            lib().rt_renderer_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, spp: int, count_traversal: bool = False) -> None:
        fn = lib().rt_render_counted if count_traversal else lib().rt_render
        _check(fn(self._h, int(spp)))

    def synchronize(self) -> None:
        _check(lib().rt_synchronize(self._h))

    def reset(self) -> None:
        _check(lib().rt_reset(self._h))

    def reset_stats(self) -> None:
        _check(lib().rt_reset_stats(self._h))

    def read_accum(self, which: int = RT_ACCUM_MEAN) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(lib().rt_read_accum(self._h, int(which), _fptr(out)))
        return out

    def read_rgba8(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        _check(lib().rt_read_rgba8(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def read_gather(self, mode: int) -> np.ndarray:
        """GatherKernel view (RT_GATHER_*): (H, W, 4) uint8, top row first."""
        out = np.zeros((self.height, self.width, 4), np.uint8)
        _check(lib().rt_read_gather(self._h, mode, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return out

    def save_image(self, which: int, path: str) -> None:
        """Write RT_IMAGE_PNG_COLOR / _PNG_DENOISED (8-bit PNG) or RT_IMAGE_PFM_MEAN / _PFM_DENOISED."""
        _check(lib().rt_save_image(self._h, which, os.fsencode(path)))

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().rt_stats(self._h, ctypes.byref(s)))
        return s

    # ---- real-time frames
    def camera(self) -> Camera:
        c = Camera()
        _check(lib().rt_get_camera(self._h, ctypes.byref(c)))
        return c

    def set_camera(self, cam: Camera) -> None:
        _check(lib().rt_set_camera(self._h, ctypes.byref(cam)))

    def render_frame(self, spp: int = 1, params: Optional[DenoiseParams] = None) -> None:
        _check(lib().rt_render_frame(self._h, int(spp), ctypes.byref(params) if params is not None else None))

    def reset_history(self) -> None:
        _check(lib().rt_reset_history(self._h))

    def read_frame(self, which: int = RT_FRAME_DENOISED) -> np.ndarray:
        dt, ch = FRAME_LAYOUT.get(which, (np.float32, 4))
        out = np.zeros((self.height, self.width, ch), dt)
        _check(lib().rt_read_frame(self._h, int(which), out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out


class Job:
    """A render job for a Pool: an accumulate job (frames = 0, `spp` samples in
    total) or a real-time job (`frames` > 0 frames of `spp` samples each, one
    camera per frame, SVGF-denoised)."""

    def __init__(self, scene: int, spp: int = 1, frames: int = 0, cameras=None,
                 render: Optional[dict] = None, denoise: Optional[DenoiseParams] = None):
        self.scene, self.spp, self.frames = scene, spp, frames
        self.cameras = list(cameras) if cameras is not None else None
        self.render, self.denoise = render, denoise

    def to_c(self, keep: list) -> JobDescC:
        d = JobDescC()
        d.scene, d.spp, d.frames = self.scene, self.spp, self.frames
        if self.cameras:
            arr = (Camera * len(self.cameras))(*self.cameras)
            keep.append(arr)
            d.cameras = ctypes.cast(arr, ctypes.POINTER(Camera))
        if self.render is not None:
            rp = RenderParams()
            lib().rt_render_params_default(ctypes.byref(rp))
            for k, v in self.render.items():
                setattr(rp, k, v)
            keep.append(rp)
            d.render = ctypes.pointer(rp)
        if self.denoise is not None:
            keep.append(self.denoise)
            d.denoise = ctypes.pointer(self.denoise)
        return d


class Pool:
    """Many jobs over resident scenes, served by one wavefront engine."""

    def __init__(self, queue_budget_bytes: int = 0, hbm_budget_bytes: int = 0, profile: bool = False,
                 pipelined: bool = False, max_realtime_jobs: int = 0):
        p = PoolParams(queue_budget_bytes, hbm_budget_bytes, int(profile), int(pipelined), int(max_realtime_jobs))
        self._h = ctypes.c_void_p()
        _check(lib().rt_pool_create(ctypes.byref(p), ctypes.byref(self._h)))
        self._scenes = []

    def close(self) -> None:
        if self._h:
            lib().rt_pool_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: Scene) -> int:
        slot = ctypes.c_int()
        _check(lib().rt_pool_add_scene(self._h, scene.handle, ctypes.byref(slot)))
        self._scenes.append(scene)  # the scene must outlive the pool
        return slot.value

    def submit(self, job: Job) -> int:
        """Job id; raises RtError(RT_ERR_BUSY) when the HBM budget cannot admit it now."""
        keep: list = []
        d = job.to_c(keep)
        out = ctypes.c_int()
        _check(lib().rt_pool_submit(self._h, ctypes.byref(d), ctypes.byref(out)))
        return out.value

    def step(self) -> int:
        n = ctypes.c_int()
        _check(lib().rt_pool_step(self._h, ctypes.byref(n)))
        return n.value

    def run(self) -> None:
        while self.step() > 0:
            pass

    def status(self, job: int) -> JobStatus:
        st = JobStatus()
        _check(lib().rt_pool_job_status(self._h, job, ctypes.byref(st)))
        return st

    def read_accum(self, job: int, which: int, width: int, height: int) -> np.ndarray:
        out = np.zeros((height, width, 4), np.float32)
        _check(lib().rt_pool_read_accum(self._h, job, which, _fptr(out)))
        return out

    def read_frame(self, job: int, which: int, width: int, height: int) -> np.ndarray:
        dt, ch = FRAME_LAYOUT.get(which, (np.float32, 4))
        out = np.zeros((height, width, ch), dt)
        _check(lib().rt_pool_read_frame(self._h, job, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def release(self, job: int) -> None:
        _check(lib().rt_pool_release(self._h, job))

    def stats(self) -> Stats:
        s = Stats()
        _check(lib().rt_pool_stats(self._h, ctypes.byref(s)))
        return s

    def reset_stats(self) -> None:
        _check(lib().rt_pool_reset_stats(self._h))

    def tuning(self) -> PoolTuning:
        t = PoolTuning()
        _check(lib().rt_pool_get_tuning(self._h, ctypes.byref(t)))
        return t

    def set_tuning(self, grid_scale: float = 1.0, pass_paths: int = 0, tile_width: int = 16,
                   ray_sort: bool = False) -> None:
        t = PoolTuning(grid_scale, pass_paths, tile_width, int(ray_sort))
        _check(lib().rt_pool_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, sample_jobs) -> dict:
        keep: list = []
        arr = (JobDescC * len(sample_jobs))(*[j.to_c(keep) for j in sample_jobs])
        rep = TuneReport()
        _check(lib().rt_pool_autotune(self._h, arr, len(sample_jobs), ctypes.byref(rep)))
        return rep.as_dict()

    def serve(self, requests, batched=True, results=None):
        """requests: list of (arrival_s, Job).  batched: True (continuous batching), False
        (the plain loop, a fresh renderer per job) or "resident" (the plain loop on a
        renderer kept per scene).  results: optional list, one float32 array per request
        (or None) of the job's image shape (h, w, 4), filled with its result when it
        completes (accumulate: mean colour; real-time: last denoised frame).  Returns
        (report dict, per-request frame latencies in ms, per-request job latencies in ms)."""
        mode = 2 if batched == "resident" else (1 if batched else 0)
        keep: list = []
        n = len(requests)
        arr = (ServeRequestC * n)()
        nframes = []
        for k, (t, job) in enumerate(requests):
            arr[k].arrival_s = t
            arr[k].job = job.to_c(keep)
            if results is not None and results[k] is not None:
                buf = results[k]
                assert buf.dtype == np.float32 and buf.flags.c_contiguous
                arr[k].result_rgba = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            nframes.append(max(1, job.frames))
        fl = np.zeros(sum(nframes), np.float64)
        jl = np.zeros(n, np.float64)
        rep = ServeReport()
        _check(lib().rt_pool_serve(self._h, arr, n, mode, ctypes.byref(rep),
                                   fl.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   jl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        per_req, o = [], 0
        for c in nframes:
            per_req.append(fl[o:o + c].copy())
            o += c
        return rep.as_dict(), per_req, jl


def device_math(op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros_like(a)
    _check(lib().rt_device_math(int(op), _fptr(a), _fptr(b), _fptr(out), a.size))
    return out
