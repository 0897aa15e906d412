"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Used by tests/, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s cpu_baseline leg.  The scene is read here independently of the
library's loader: JSON semantics follow the reference's Scene::Parse*
(Src/Core/scene.cpp:187-616); OBJ import and object transforms are
rtref_obj's restatement of OBJLoader::LoadObject (Src/Utils/objloader.cpp:66-210),
tinyobjloader and Eigen.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import rtref_obj
from .rtref_image import read_texture

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "librtref.so")
MAT_STRIDE = 24
MAT_TYPES = {"light": 0, "diffuse": 1, "pbr": 2, "dielectric": 3, "glass": 3, "subsurface": 4}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class Desc(ctypes.Structure):
    _fields_ = [
        ("num_tris", ctypes.c_int), ("pos", ctypes.c_void_p), ("nrm", ctypes.c_void_p), ("uv", ctypes.c_void_p),
        ("mesh", ctypes.c_void_p), ("num_meshes", ctypes.c_int), ("mesh_material", ctypes.c_void_p),
        ("mesh_light", ctypes.c_void_p), ("num_materials", ctypes.c_int), ("materials", ctypes.c_void_p),
        ("num_lights", ctypes.c_int), ("light_type", ctypes.c_void_p), ("light_mesh", ctypes.c_void_p),
        ("light_emission", ctypes.c_void_p), ("env_w", ctypes.c_int), ("env_h", ctypes.c_int),
        ("env_rgb", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int), ("fovy", ctypes.c_float),
        ("direction", ctypes.c_float * 3), ("num_textures", ctypes.c_int), ("tex_info", ctypes.c_void_p),
        ("tex_rgba", ctypes.c_void_p),
    ]


class Params(ctypes.Structure):
    _fields_ = [("max_bounces", ctypes.c_int), ("rr", ctypes.c_float), ("rr_start_depth", ctypes.c_int),
                ("seed", ctypes.c_uint32), ("flush_survivors", ctypes.c_int), ("env_alias", ctypes.c_int),
                ("shadow_mode", ctypes.c_int)]


class Orbit(ctypes.Structure):
    _fields_ = [("target", ctypes.c_float * 3), ("distance", ctypes.c_float), ("theta_deg", ctypes.c_float),
                ("phi_deg", ctypes.c_float)]


class GBufPtrs(ctypes.Structure):
    _fields_ = [("pos_tri", ctypes.c_void_p), ("nrm_mat", ctypes.c_void_p), ("bary_mesh", ctypes.c_void_p),
                ("motion_depth", ctypes.c_void_p)]


class SvgfParams(ctypes.Structure):
    """DenoiseParam (Src/Utils/ui.h:14-25) with its defaults."""
    _fields_ = [("max_history", ctypes.c_int), ("sigma_light", ctypes.c_float), ("sigma_normal", ctypes.c_float),
                ("sigma_depth", ctypes.c_float), ("reproj_depth", ctypes.c_float), ("atrous_iters", ctypes.c_int)]

    @classmethod
    def default(cls, **kw) -> "SvgfParams":
        p = cls(24, 8.0, 256.0, 1.0, 0.1, 5)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ref_scene_create.argtypes = [ctypes.POINTER(Desc)]
        L.ref_scene_create.restype = vp
        L.ref_scene_destroy.argtypes = [vp]
        L.ref_scene_refit.argtypes = [vp, vp, vp]
        L.ref_scene_refit.restype = None
        L.ref_bvh_num_nodes.argtypes = [vp]
        L.ref_bvh_max_depth.argtypes = [vp]
        L.ref_bvh_nodes.argtypes = [vp, vp]
        L.ref_bvh_order.argtypes = [vp, vp]
        L.ref_camera.argtypes = [vp, vp]
        L.ref_trace.argtypes = [vp, vp, ctypes.c_int, vp, vp]
        L.ref_occluded.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp, vp]
        L.ref_render.argtypes = [vp, ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp, vp]
        L.ref_render_f64.argtypes = [vp, ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.ref_philox.argtypes = [vp, vp, vp]
        L.ref_math.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.ref_math.restype = ctypes.c_float
        L.ref_bsdf_eval.argtypes = [vp, vp, vp, vp, vp]
        L.ref_bsdf_eval.restype = ctypes.c_float
        L.ref_bsdf_sample.argtypes = [vp, vp, vp, vp, vp]
        L.ref_bsdf_sample.restype = ctypes.c_float
        L.ref_fresnel_dielectric.argtypes = [ctypes.c_float] * 3
        L.ref_fresnel_dielectric.restype = ctypes.c_float
        L.ref_ggx_d.argtypes = [vp, vp, vp, vp]
        L.ref_ggx_d.restype = ctypes.c_float
        L.ref_burley.argtypes = [vp, ctypes.c_float, ctypes.c_int]
        L.ref_burley.restype = ctypes.c_float
        L.ref_sss_dist.argtypes = [ctypes.c_float, ctypes.c_float]
        L.ref_sss_dist.restype = ctypes.c_float
        L.ref_sss_probe.argtypes = [vp, vp, vp, ctypes.c_float, ctypes.c_float, vp, vp]
        L.ref_sss_walk.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp,
                                   ctypes.POINTER(ctypes.c_int)]
        L.ref_sss_walk.restype = ctypes.c_int
        L.ref_sss_walk_batch.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, vp, ctypes.c_uint32, vp, vp, vp,
                                         vp, vp]
        L.ref_env_tables.argtypes = [vp, vp, vp]
        L.ref_env_alias_tables.argtypes = [vp, vp, vp]
        L.ref_env_sample_alias.argtypes = [vp, ctypes.c_float, ctypes.c_float, vp]
        L.ref_alias_build.argtypes = [vp, ctypes.c_int, vp, vp]
        L.ref_env_sample.argtypes = [vp, ctypes.c_float, ctypes.c_float, vp]
        L.ref_env_pdf_uv.argtypes = [vp, ctypes.c_float, ctypes.c_float]
        L.ref_env_pdf_uv.restype = ctypes.c_float
        L.ref_texture_albedo.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, vp]
        L.ref_texture_albedo.restype = None
        L.ref_get_orbit.argtypes = [vp, ctypes.POINTER(Orbit)]
        L.ref_set_orbit.argtypes = [vp, ctypes.POINTER(Orbit)]
        L.ref_view_projection.argtypes = [vp, ctypes.POINTER(Orbit), ctypes.c_float, ctypes.c_float, vp]
        L.ref_gbuffer.argtypes = [vp, ctypes.POINTER(Orbit), ctypes.POINTER(GBufPtrs)]
        L.ref_svgf_temporal.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SvgfParams),
                                        ctypes.c_float] + [vp] * 7 + [ctypes.POINTER(GBufPtrs)] * 2 + [vp] * 5
        L.ref_svgf_variance.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(SvgfParams)] + [vp] * 5 + \
            [ctypes.POINTER(GBufPtrs), vp, vp]
        L.ref_svgf_atrous.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(SvgfParams)] + \
            [vp] * 3 + [ctypes.POINTER(GBufPtrs), vp, vp]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ scene parse
def _f32(x) -> np.float32:
    return np.float32(x)


def _read_env(path: str):
    with open(path, "rb") as f:
        data = f.read()
    if data[:2] in (b"PF", b"Pf"):
        color = data[:2] == b"PF"
        toks, p = [], 2
        while len(toks) < 3:
            while data[p:p + 1].isspace():
                p += 1
            b = p
            while not data[p:p + 1].isspace():
                p += 1
            toks.append(data[b:p])
        w, h, scale = int(toks[0]), int(toks[1]), float(toks[2])
        p += 1
        ch = 3 if color else 1
        arr = np.frombuffer(data, dtype="<f4" if scale < 0 else ">f4", count=w * h * ch, offset=p)
        arr = arr.reshape(h, w, ch)[::-1].astype(np.float32)
        if ch == 1:
            arr = np.repeat(arr, 3, axis=2)
        return w, h, np.ascontiguousarray(arr)
    return _read_rgbe(data, path)


def _read_rgbe(data: bytes, path: str):
    """Radiance RGBE as HDRTexture::Load reads it (Src/Core/texture.cu:17-25,
    stbi_loadf(..., 4)): the vendored stb_image 2.30 reader
    (ThirdParty/stb_image/stb_image.h:7083-7288) restated.  Returns (w, h, rgb)
    with row 0 = the first scanline; raises ValueError where stb fails, and on
    data that ends early or a zero size (stb's output is unspecified there)."""
    p = 0

    def get8():  # stbi__get8
        nonlocal p
        if p >= len(data):
            raise ValueError(f"truncated RGBE data: {path}")
        p += 1
        return data[p - 1]

    def token():  # stbi__hdr_gettoken (stb_image.h:7108-7128)
        nonlocal p
        t = bytearray()
        if p >= len(data):
            return bytes(t)
        c = data[p]
        p += 1
        while p < len(data) and c != 10:
            t.append(c)
            if len(t) == 1023:
                while p < len(data):
                    p += 1
                    if data[p - 1] == 10:
                        break
                break
            c = data[p]
            p += 1
        return bytes(t)

    if not (data.startswith(b"#?RADIANCE\n") or data.startswith(b"#?RGBE\n")):  # stbi__hdr_test
        raise ValueError(f"not a Radiance RGBE file: {path}")
    token()
    valid = False
    while True:  # header lines up to an empty one (7179-7183)
        t = token()
        if not t:
            break
        valid |= t == b"FORMAT=32-bit_rle_rgbe"
    if not valid:
        raise ValueError(f"unsupported RGBE format: {path}")
    res = token().decode("latin-1")
    if not res.startswith("-Y "):  # 7189-7196
        raise ValueError(f"unsupported RGBE layout: {path}")

    def strtol(s: str):
        s2 = s.lstrip(" \t\n\v\f\r")
        k = 0
        if k < len(s2) and s2[k] in "+-":
            k += 1
        while k < len(s2) and s2[k].isdigit():
            k += 1
        digits = s2[:k]
        val = int(digits) if digits not in ("", "+", "-") else 0
        return val, (s2[k:] if digits not in ("", "+", "-") else s)

    h, rest = strtol(res[3:])
    rest = rest.lstrip(" ")
    if not rest.startswith("+X "):
        raise ValueError(f"unsupported RGBE layout: {path}")
    w, _ = strtol(rest[3:])
    if h > (1 << 24) or w > (1 << 24) or h < 0 or w < 0 or (w > 0 and h > (2**31 - 1) // 16 // w):
        raise ValueError(f"RGBE size out of range: {path}")
    if w == 0 or h == 0:
        raise ValueError(f"zero-sized RGBE image: {path}")
    px = np.zeros((h * w, 4), np.uint8)

    def flat_from(texel):  # the flat loop (7208-7216)
        nonlocal p
        n = (h * w - texel) * 4
        if len(data) - p < n:
            raise ValueError(f"truncated RGBE data: {path}")
        px[texel:] = np.frombuffer(data, np.uint8, n, p).reshape(-1, 4)
        p += n

    if w < 8 or w >= 32768:
        flat_from(0)
    else:
        for y in range(h):  # run-length scanlines (7218-7282)
            c1, c2, ln = get8(), get8(), get8()
            if c1 != 2 or c2 != 2 or (ln & 0x80):
                px[0] = (c1, c2, ln, get8())  # pixel 0 of row 0, then flat from row 0, pixel 1
                flat_from(1)
                break
            if ((ln << 8) | get8()) != w:
                raise ValueError(f"invalid RGBE scanline length: {path}")
            row = np.zeros((w, 4), np.uint8)
            for c in range(4):
                x = 0
                while x < w:
                    count = get8()
                    if count > 128:
                        value = get8()
                        count -= 128
                        if count == 0 or count > w - x:
                            raise ValueError(f"bad RGBE run-length data: {path}")
                        row[x:x + count, c] = value
                    else:
                        if count == 0 or count > w - x:
                            raise ValueError(f"bad RGBE run-length data: {path}")
                        if len(data) - p < count:
                            raise ValueError(f"truncated RGBE data: {path}")
                        row[x:x + count, c] = np.frombuffer(data, np.uint8, count, p)
                        p += count
                    x += count
            px[y * w:(y + 1) * w] = row
    # stbi__hdr_convert (7131-7155): byte * 2^(e - 136), zero when e == 0.
    e = px[:, 3].astype(np.int32)
    f = np.where(e != 0, np.ldexp(np.float32(1.0), e - 136).astype(np.float32), np.float32(0.0))
    rgb = px[:, :3].astype(np.float32) * f[:, None].astype(np.float32)
    return w, h, np.ascontiguousarray(rgb.reshape(h, w, 3).astype(np.float32))


@dataclass
class SceneArrays:
    pos: np.ndarray
    nrm: np.ndarray
    uv: np.ndarray
    mesh: np.ndarray
    mesh_material: np.ndarray
    mesh_light: np.ndarray
    materials: np.ndarray
    light_type: np.ndarray
    light_mesh: np.ndarray
    light_emission: np.ndarray
    env_w: int
    env_h: int
    env_rgb: Optional[np.ndarray]
    width: int
    height: int
    fovy: float
    direction: tuple
    max_bounces: int
    rr: float
    spp: int
    seed: int
    rr_start_depth: int
    num_lights: int = 0
    textures: list = field(default_factory=list)  # (h, w, 4) uint8 per "textures" entry


def parse_scene(json_path: str, geometry=None) -> SceneArrays:
    """The scene as Scene::ParseSceneFile reads it.  geometry: optional, one
    list of meshes per entry of "objects" ({name, material, pos, nrm, uv}, as
    rtref_obj.load_obj returns them) used instead of reading the OBJ files, for
    scenes too large for this module's Python OBJ reader (the reader itself is
    checked against the library's importer and the reference's libraries on
    smaller files)."""
    base = os.path.dirname(os.path.abspath(json_path))
    with open(json_path, encoding="utf-8-sig") as f:  # a leading BOM is skipped, as nlohmann does
        doc = json.load(f)
    cam = doc["camera"]
    W, H = (cam["resolution"] if "resolution" in cam else (800, 600))
    mats_json = doc["materials"]
    materials = np.zeros((max(1, len(mats_json)), MAT_STRIDE), np.float32)
    materials[:, 0] = 1
    materials[:, 7] = 0.5
    materials[:, 9] = 1.0
    materials[:, 10] = 1.0
    materials[:, 19:22] = 1.0  # sigma_s default
    # Textures (scene.cpp:618-644): every entry decoded in order; an entry
    # without 'path' or 'name' keeps an empty slot.
    textures, tex_index = [], {}
    for t in doc["textures"]:
        if "path" not in t or "name" not in t:
            textures.append(np.zeros((0, 0, 4), np.uint8))
            continue
        tp = t["path"] if os.path.isabs(t["path"]) else os.path.join(base, t["path"])
        tex_index[t["name"]] = len(textures)
        textures.append(read_texture(tp))
    name_to_mat = {}
    for i, m in enumerate(mats_json):
        if "name" not in m or "type" not in m or m["type"] not in MAT_TYPES:
            continue
        t = MAT_TYPES[m["type"]]
        r = materials[i]
        r[0] = t
        if t == 1:
            r[1:4] = m.get("basecolor", [0, 0, 0])
        elif t == 0:
            r[4:7] = m.get("emission", [0, 0, 0])
        elif t == 2:
            r[1:4] = m.get("basecolor", [0, 0, 0])
            r[8] = m.get("metallic", 0.0)
            r[7] = m.get("roughness", 0.5)
        elif t == 4:
            r[1:4] = m.get("basecolor", [0, 0, 0])
            r[11:14] = m.get("scatter_distance", [1, 1, 1])
            if "ior" in m:
                r[10] = m["ior"]
            if "roughness" in m:
                r[7] = m["roughness"]
            r[8] = 0
            # The library's random-walk mode (absent from the reference's schema).
            method = m.get("subsurface_method", "burley")
            if method not in ("burley", "random_walk"):
                raise ValueError(f"unsupported subsurface_method {method!r}")
            r[22] = 1 if method == "random_walk" else 0
            r[16:19] = m.get("sigma_a", [0, 0, 0])
            r[19:22] = m.get("sigma_s", [1, 1, 1])
        elif t == 3:
            r[1:4] = m.get("basecolor", [1, 1, 1])
            r[10] = m.get("ior", 1.5)
            r[9] = m.get("transmission", 1.0)
            r[7] = m.get("roughness", 0.0)
            r[8] = 0
        if m.get("diffuse_texture") in tex_index:
            r[14] = tex_index[m["diffuse_texture"]] + 1
        name_to_mat[m["name"]] = i
    pos, nrm, uv, mesh, mmat, mnames = [], [], [], [], [], []
    for oi, o in enumerate(doc["objects"]):
        if "name" not in o or "path" not in o:
            continue
        if geometry is not None:
            loaded = geometry[oi]
        else:
            xf, xf_inv = rtref_obj.object_transform(o)
            path = o["path"] if os.path.isabs(o["path"]) else os.path.join(base, o["path"])
            loaded = rtref_obj.load_obj(path, xf, xf_inv)
        expanded = len(loaded) > 1
        for lm in loaded:
            if "material" in o:
                mn = o["material"]
            elif isinstance(o.get("material_overrides"), dict) and lm["material"] in o["material_overrides"]:
                mn = o["material_overrides"][lm["material"]]
            elif lm["material"] and lm["material"] in name_to_mat:
                mn = lm["material"]
            else:
                mn = o.get("default_material", "")
            if mn not in name_to_mat:
                continue
            mid = len(mmat)
            mmat.append(name_to_mat[mn])
            mnames.append(o["name"] + "::" + lm["name"] if expanded else o["name"])
            pos.append(lm["pos"])
            nrm.append(lm["nrm"])
            uv.append(lm["uv"])
            mesh.append(np.full(len(lm["pos"]), mid, np.int32))
    mesh_light = np.full(max(1, len(mmat)), -1, np.int32)
    lj = doc["lights"]
    lt = np.zeros(max(1, len(lj)), np.int32)
    lmsh = np.full(max(1, len(lj)), -1, np.int32)
    lem = np.zeros((max(1, len(lj)), 3), np.float32)
    env = (0, 0, None)
    for i, L in enumerate(lj):
        if L.get("type") == "area":
            if "ke" in L:
                lem[i] = L["ke"]
            if "material" in L and L["material"] in name_to_mat:
                lem[i] = materials[name_to_mat[L["material"]], 4:7]
            m = mnames.index(L["lightMeshName"])
            lmsh[i] = m
            mesh_light[m] = i
        elif L.get("type") == "env":
            lt[i] = 2
            p = L["path"] if os.path.isabs(L["path"]) else os.path.join(base, L["path"])
            env = _read_env(p)
    return SceneArrays(
        pos=np.ascontiguousarray(np.concatenate(pos or [np.zeros((0, 9), np.float32)]).astype(np.float32)),
        nrm=np.ascontiguousarray(np.concatenate(nrm or [np.zeros((0, 9), np.float32)]).astype(np.float32)),
        uv=np.ascontiguousarray(np.concatenate(uv or [np.zeros((0, 6), np.float32)]).astype(np.float32)),
        mesh=np.concatenate(mesh or [np.zeros(0, np.int32)]).astype(np.int32), mesh_material=np.array(mmat if mmat else [0], np.int32),
        mesh_light=mesh_light, materials=materials, light_type=lt, light_mesh=lmsh, light_emission=lem,
        env_w=env[0], env_h=env[1], env_rgb=env[2], width=int(W), height=int(H), fovy=float(cam.get("fov", 45.0)),
        direction=tuple(cam.get("direction", [0, 0, -1])), max_bounces=int(cam.get("max_bounces", 5)),
        rr=float(cam.get("rr", 0.8)), spp=int(cam.get("spp", 16)), seed=int(cam.get("seed", 1234)),
        rr_start_depth=int(cam.get("rr_start_depth", 0)), num_lights=len(lj), textures=textures,
    )


def geometry_from_arrays(doc_objects: list, pos, nrm, uv, mesh_start, mesh_count):
    """`geometry` for parse_scene from scene-order triangle arrays holding one
    mesh per object, in object order (a loader's output for scenes whose
    objects each import as a single mesh)."""
    if len(mesh_start) != len(doc_objects):
        raise ValueError("geometry_from_arrays needs exactly one mesh per object")
    out = []
    for o, s0, n in zip(doc_objects, mesh_start, mesh_count):
        sl = slice(int(s0), int(s0) + int(n))
        out.append([{"name": o.get("name", ""), "material": "", "pos": pos[sl], "nrm": nrm[sl], "uv": uv[sl]}])
    return out


class RefScene:
    """The oracle's scene: BVH, camera and env tables built in C from SceneArrays."""

    def __init__(self, a: SceneArrays):
        self.a = a
        d = Desc()
        d.num_tris = a.pos.shape[0]
        d.pos, d.nrm, d.uv, d.mesh = _p(a.pos), _p(a.nrm), _p(a.uv), _p(a.mesh)
        d.num_meshes = int(a.mesh_material.shape[0])
        d.mesh_material, d.mesh_light = _p(a.mesh_material), _p(a.mesh_light)
        d.num_materials = int(a.materials.shape[0])
        d.materials = _p(a.materials)
        d.num_lights = int(a.num_lights)
        d.light_type, d.light_mesh, d.light_emission = _p(a.light_type), _p(a.light_mesh), _p(a.light_emission)
        d.env_w, d.env_h = a.env_w, a.env_h
        d.env_rgb = _p(a.env_rgb) if a.env_rgb is not None else None
        d.width, d.height, d.fovy = a.width, a.height, a.fovy
        d.direction[:] = [np.float32(x) for x in a.direction]
        if a.textures:
            info, first = [], 0
            for t in a.textures:
                info += [first, t.shape[1], t.shape[0]]
                first += t.shape[0] * t.shape[1]
            self._tex_info = np.array(info, np.int32)
            self._tex_rgba = np.ascontiguousarray(
                np.concatenate([t.reshape(-1) for t in a.textures] + [np.zeros(4, np.uint8)]))
            d.num_textures = len(a.textures)
            d.tex_info, d.tex_rgba = _p(self._tex_info), _p(self._tex_rgba)
        self._desc = d
        self._h = lib().ref_scene_create(ctypes.byref(d))

    def __del__(self):
        try:
            if self._h:
                lib().ref_scene_destroy(self._h)
        except Exception:
            pass

    def params(self, max_bounces=None, rr=None, rr_start_depth=None, seed=None, flush_survivors=False,
               env_alias=False, shadow_mode=0) -> Params:
        a = self.a
        p = Params()
        p.max_bounces = a.max_bounces if max_bounces is None else max_bounces
        p.rr = a.rr if rr is None else rr
        p.rr_start_depth = a.rr_start_depth if rr_start_depth is None else rr_start_depth
        p.seed = a.seed if seed is None else seed
        p.flush_survivors = int(flush_survivors)
        p.env_alias = int(env_alias)
        p.shadow_mode = int(shadow_mode)
        return p

    def bvh(self):
        n = lib().ref_bvh_num_nodes(self._h)
        nodes = np.zeros((n, 8), np.int32)
        order = np.zeros(self.a.pos.shape[0], np.int32)
        lib().ref_bvh_nodes(self._h, _p(nodes))
        lib().ref_bvh_order(self._h, _p(order))
        return nodes, order

    def refit(self, pos: np.ndarray, nrm: Optional[np.ndarray] = None) -> None:
        """New vertex positions ((n, 9) float32, scene order) and optionally shading
        normals for the same triangles; the tree's boxes are refit (ref_scene_refit)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 9)
        assert pos.shape[0] == self.a.pos.shape[0]
        self._refit_pos = pos
        self._refit_nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32).reshape(-1, 9)
        lib().ref_scene_refit(self._h, _p(pos), None if nrm is None else _p(self._refit_nrm))

    def max_depth(self) -> int:
        return lib().ref_bvh_max_depth(self._h)

    def camera(self) -> np.ndarray:
        out = np.zeros(12, np.float32)
        lib().ref_camera(self._h, _p(out))
        return out.reshape(4, 3)

    def trace(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        counts = np.zeros((n, 2), np.uint32)
        lib().ref_trace(self._h, _p(rays), n, _p(hits), _p(counts))
        return hits[:, 0].copy(), hits[:, 1].copy(), hits[:, 2].copy(), hits[:, 3].copy().view(np.int32), counts

    def occluded(self, rays: np.ndarray, brute: bool = False):
        """The library's shadow segment test (mode 2) for (n, 8) rays o.xyz, tmin, d.xyz, tmax:
        (occluded bool array, (n, 2) node visits / triangle tests); brute tests every triangle
        and counts nothing."""
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        out = np.zeros(n, np.int32)
        counts = np.zeros((n, 2), np.uint32)
        lib().ref_occluded(self._h, _p(rays), n, int(brute), _p(out), _p(counts))
        return out.astype(bool), counts

    def render(self, spp: int, params: Params, sample0: int = 0, pixels: Optional[np.ndarray] = None):
        """Returns (direct, indirect) sums, (npix, 4) float32, and [ext, shadow, subsurface]
        ray counts (subsurface: Burley probe rays and random-walk segments)."""
        npix = self.a.width * self.a.height if pixels is None else len(pixels)
        d = np.zeros((npix, 4), np.float32)
        i = np.zeros((npix, 4), np.float32)
        st = np.zeros(3, np.uint64)
        px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
        lib().ref_render(self._h, ctypes.byref(params), sample0, spp, _p(px) if px is not None else None, npix,
                         _p(d), _p(i), _p(st))
        return d, i, st

    def render_f64(self, spp: int, params: Params, sample0: int = 0, pixels: Optional[np.ndarray] = None):
        npix = self.a.width * self.a.height if pixels is None else len(pixels)
        d = np.zeros((npix, 4), np.float64)
        i = np.zeros((npix, 4), np.float64)
        px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
        lib().ref_render_f64(self._h, ctypes.byref(params), sample0, spp, _p(px) if px is not None else None, npix,
                             _p(d), _p(i))
        return d, i

    def orbit(self) -> Orbit:
        o = Orbit()
        lib().ref_get_orbit(self._h, ctypes.byref(o))
        return o

    def set_orbit(self, o: Orbit) -> None:
        lib().ref_set_orbit(self._h, ctypes.byref(o))

    def view_projection(self, o: Orbit, near: float, far: float) -> np.ndarray:
        out = np.zeros(16, np.float32)
        lib().ref_view_projection(self._h, ctypes.byref(o), near, far, _p(out))
        return out.reshape(4, 4)

    def gbuffer(self, prev: Optional[Orbit] = None) -> "GBuffer":
        g = GBuffer.empty(self.a.width, self.a.height)
        lib().ref_gbuffer(self._h, ctypes.byref(prev if prev is not None else self.orbit()), ctypes.byref(g.ptrs()))
        return g

    def sss_probe(self, p, n, dist: float, angle: float):
        pp = np.ascontiguousarray(p, np.float32)
        nn = np.ascontiguousarray(n, np.float32)
        op = np.zeros(3, np.float32)
        on = np.zeros(3, np.float32)
        lib().ref_sss_probe(self._h, _p(pp), _p(nn), dist, angle, _p(op), _p(on))
        return op, on

    def sss_walk(self, mesh: int, p, n_out, e1: float, e2: float, seed: int = 1, pixel: int = 0, sample: int = 0,
                 bounce: int = 0):
        """One random walk into mesh `mesh` (a random-walk medium): (ok, exit point,
        exit normal, per-channel weight, segments traced)."""
        pp = np.ascontiguousarray(p, np.float32)
        nn = np.ascontiguousarray(n_out, np.float32)
        op, on, w = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
        st = ctypes.c_int(0)
        ok = lib().ref_sss_walk(self._h, mesh, _p(pp), _p(nn), e1, e2, seed, pixel, sample, bounce, _p(op), _p(on),
                                _p(w), ctypes.byref(st))
        return bool(ok), op, on, w, st.value

    def sss_walks(self, mesh: int, p, n_out, e, seed: int = 1):
        """len(e) walks from one entry point (walk i: pixel i, entry uniforms e[i]):
        (ok, exit points, exit normals, weights, segments traced) as arrays."""
        e = np.ascontiguousarray(e, np.float32).reshape(-1, 2)
        n = e.shape[0]
        pp = np.ascontiguousarray(p, np.float32)
        nn = np.ascontiguousarray(n_out, np.float32)
        ok, steps = np.zeros(n, np.int32), np.zeros(n, np.int32)
        op, on, w = (np.zeros((n, 3), np.float32) for _ in range(3))
        lib().ref_sss_walk_batch(self._h, mesh, _p(pp), _p(nn), n, _p(e), seed, _p(ok), _p(op), _p(on), _p(w),
                                 _p(steps))
        return ok.astype(bool), op, on, w, steps

    def env_sample(self, rv: float, ru: float):
        uv = np.zeros(2, np.float32)
        lib().ref_env_sample(self._h, rv, ru, _p(uv))
        return uv

    def texture_albedo(self, tex: int, u: float, v: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        lib().ref_texture_albedo(self._h, tex, u, v, _p(out))
        return out

    def env_pdf_uv(self, u: float, v: float) -> float:
        return float(lib().ref_env_pdf_uv(self._h, u, v))

    def env_sample_alias(self, rv: float, ru: float):
        uv = np.zeros(2, np.float32)
        lib().ref_env_sample_alias(self._h, rv, ru, _p(uv))
        return uv

    def env_alias_tables(self):
        """(row thresholds (h, w), row aliases (h, w), marginal thresholds (h,), marginal aliases (h,))."""
        h, w = self.a.env_h, self.a.env_w
        row = np.zeros((h, w, 2), np.float32)
        marg = np.zeros((h, 2), np.float32)
        lib().ref_env_alias_tables(self._h, _p(row), _p(marg))
        return row[..., 0], row[..., 1].copy().view(np.int32), marg[:, 0], marg[:, 1].copy().view(np.int32)

    def env_tables(self):
        row = np.zeros((self.a.env_h, self.a.env_w), np.float32)
        marg = np.zeros(self.a.env_h, np.float32)
        lib().ref_env_tables(self._h, _p(row), _p(marg))
        return row, marg


def philox(ctr, key) -> np.ndarray:
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().ref_philox(_p(c), _p(k), _p(out))
    return out


def math(op: int, a: float, b: float = 0.0) -> float:
    return float(lib().ref_math(op, a, b))


# ------------------------------------------------------------------ SVGF
@dataclass
class GBuffer:
    """G-buffer planes, (height, width, 4) each, row 0 = bottom."""
    pos_tri: np.ndarray
    nrm_mat: np.ndarray
    bary_mesh: np.ndarray
    motion_depth: np.ndarray

    @classmethod
    def empty(cls, w: int, h: int) -> "GBuffer":
        return cls(np.zeros((h, w, 4), np.float32), np.zeros((h, w, 4), np.uint16), np.zeros((h, w, 4), np.uint16),
                   np.zeros((h, w, 4), np.float32))

    def ptrs(self) -> GBufPtrs:
        for a in (self.pos_tri, self.nrm_mat, self.bary_mesh, self.motion_depth):
            assert a.flags.c_contiguous
        return GBufPtrs(_p(self.pos_tri), _p(self.nrm_mat), _p(self.bary_mesh), _p(self.motion_depth))


def _c(a: np.ndarray, dt) -> np.ndarray:
    a = np.ascontiguousarray(a, dt)
    return a


def svgf_temporal(first: bool, p: SvgfParams, inv_spp, acc_d, acc_i, hist_prev_d, hist_prev_i, mom_prev_d,
                  mom_prev_i, len_prev, cur: GBuffer, prev: GBuffer):
    """TemporalDenoising (renderer.cuh:524-677): returns (hist_d, hist_i, mom_d, mom_i, len)."""
    h, w = acc_d.shape[:2]
    acc_d, acc_i = _c(acc_d, np.float32), _c(acc_i, np.float32)
    hpd, hpi = _c(hist_prev_d, np.float32), _c(hist_prev_i, np.float32)
    mpd, mpi = _c(mom_prev_d, np.float32), _c(mom_prev_i, np.float32)
    lp = _c(len_prev, np.int32)
    hd, hi = np.zeros((h, w, 4), np.float32), np.zeros((h, w, 4), np.float32)
    md, mi = np.zeros((h, w, 2), np.float32), np.zeros((h, w, 2), np.float32)
    ln = np.zeros((h, w), np.int32)
    cp, pp = cur.ptrs(), prev.ptrs()
    lib().ref_svgf_temporal(w, h, int(first), ctypes.byref(p), float(np.float32(inv_spp)), _p(acc_d), _p(acc_i),
                            _p(hpd), _p(hpi), _p(mpd), _p(mpi), _p(lp), ctypes.byref(cp), ctypes.byref(pp),
                            _p(hd), _p(hi), _p(md), _p(mi), _p(ln))
    return hd, hi, md, mi, ln


def svgf_variance(p: SvgfParams, in_d, in_i, mom_d, mom_i, length, g: GBuffer):
    """ComputeVarianceKernel (renderer.cuh:689-805): returns (out_d, out_i)."""
    h, w = in_d.shape[:2]
    in_d, in_i = _c(in_d, np.float32), _c(in_i, np.float32)
    mom_d, mom_i, length = _c(mom_d, np.float32), _c(mom_i, np.float32), _c(length, np.int32)
    od, oi = np.zeros((h, w, 4), np.float32), np.zeros((h, w, 4), np.float32)
    gp = g.ptrs()
    lib().ref_svgf_variance(w, h, ctypes.byref(p), _p(in_d), _p(in_i), _p(mom_d), _p(mom_i), _p(length),
                            ctypes.byref(gp), _p(od), _p(oi))
    return od, oi


def svgf_atrous(step: int, p: SvgfParams, in_d, in_i, length, g: GBuffer):
    """AtrousKernel (renderer.cuh:807-954): returns (out_d, out_i)."""
    h, w = in_d.shape[:2]
    in_d, in_i, length = _c(in_d, np.float32), _c(in_i, np.float32), _c(length, np.int32)
    od, oi = np.zeros((h, w, 4), np.float32), np.zeros((h, w, 4), np.float32)
    gp = g.ptrs()
    lib().ref_svgf_atrous(w, h, int(step), ctypes.byref(p), _p(in_d), _p(in_i), _p(length), ctypes.byref(gp),
                          _p(od), _p(oi))
    return od, oi


class RefFrames:
    """The oracle's real-time loop (RenderLoop with denoising, renderer.cuh:1141-1302):
    per frame the G-buffer against the previous frame's camera, `spp` fresh samples
    (sample indices frame*spp ...), temporal blend, variance, and the a-trous passes
    (none on the first frame after a history reset)."""

    def __init__(self, scene: RefScene, params: Params, svgf: Optional[SvgfParams] = None):
        self.s, self.params = scene, params
        self.p = svgf if svgf is not None else SvgfParams.default()
        w, h = scene.a.width, scene.a.height
        self.w, self.h = w, h
        self.frames = 0
        self.last_orbit: Optional[Orbit] = None
        z4 = lambda: np.zeros((h, w, 4), np.float32)
        self.hist = (z4(), z4())
        self.mom = (np.zeros((h, w, 2), np.float32), np.zeros((h, w, 2), np.float32))
        self.length = np.zeros((h, w), np.int32)
        self.g: Optional[GBuffer] = None

    def reset_history(self) -> None:
        self.frames = 0

    def frame(self, spp: int = 1, orbit: Optional[Orbit] = None) -> dict:
        s, w, h = self.s, self.w, self.h
        if orbit is not None:
            s.set_orbit(orbit)
        cur = s.orbit()
        first = self.frames == 0
        prev_orbit = cur if first else self.last_orbit
        g = s.gbuffer(prev_orbit)
        d, i, _ = s.render(spp, self.params, sample0=self.frames * spp)
        acc_d, acc_i = d.reshape(h, w, 4), i.reshape(h, w, 4)
        inv = np.float32(1.0) / np.float32(spp)
        prev_g = g if self.g is None else self.g
        hd, hi, md, mi, ln = svgf_temporal(first, self.p, inv, acc_d, acc_i, self.hist[0], self.hist[1],
                                           self.mom[0], self.mom[1], self.length, g, prev_g)
        vd, vi = svgf_variance(self.p, hd, hi, md, mi, ln, g)
        od, oi = vd, vi
        if not first:
            for k in range(max(1, min(8, self.p.atrous_iters))):
                od, oi = svgf_atrous(1 << k, self.p, od, oi, ln, g)
        den = np.empty((h, w, 4), np.float32)
        den[..., :3] = od[..., :3] + oi[..., :3]
        den[..., 3] = 1.0
        self.hist, self.mom, self.length, self.g = (hd, hi), (md, mi), ln, g
        self.last_orbit = cur
        self.frames += 1
        return {"acc_d": acc_d, "acc_i": acc_i, "gbuffer": g, "hist_d": hd, "hist_i": hi, "mom_d": md,
                "mom_i": mi, "length": ln, "var_d": vd, "var_i": vi, "out_d": od, "out_i": oi, "denoised": den}


def alias_build(weights):
    """Walker/Vose alias table of the weights: (thresholds float32, aliases int32)."""
    w = np.ascontiguousarray(weights, np.float64)
    prob = np.zeros(len(w), np.float32)
    alias = np.zeros(len(w), np.int32)
    lib().ref_alias_build(_p(w), len(w), _p(prob), _p(alias))
    return prob, alias
