"""Procedural scenes in the reference's scene-JSON + OBJ format.

The reference ships no assets (its scene JSONs point at absolute Windows
paths), so the benchmark workloads are generated here and written in the
same schema the reference parses (Src/Core/scene.cpp:129-644):

* ``cornell_small``   -- 5 wall quads, a short and a tall axis-aligned box and a
  quad area light (36 triangles), Lambertian albedos, 256x256, 16 spp, depth 8.
* ``cornell_spheres`` -- the same shell plus 4 icospheres (subdivision 4,
  5,120 triangles each, radius U[0.15, 0.25]) with GGX metallic-roughness
  (roughness U[0.05, 0.8], metallic in {0, 1}) and one IOR-1.5 dielectric,
  512x512, 64 spp, depth 12.
* ``textured_cornell`` -- the shell with PNG diffuse textures on diffuse, GGX,
  subsurface and dielectric materials (coordinates past [0, 1], 1x1 and 5x3
  textures, wrapped sphere longitudes), 96x96, 4 spp, depth 6.

Floats are written with 9 significant digits so the OBJ text round-trips
every float32 exactly.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

WHITE = (0.73, 0.73, 0.73)
RED = (0.65, 0.05, 0.05)
GREEN = (0.12, 0.45, 0.15)
LIGHT_RADIANCE = (17.0, 12.0, 4.0)


@dataclass
class MeshSpec:
    name: str
    material: str
    pos: Optional[np.ndarray] = None   # (n, 3, 3) float32 triangle soup
    nrm: Optional[np.ndarray] = None   # (n, 3, 3) float32 per-vertex normals or None
    verts: Optional[np.ndarray] = None  # indexed form: (m, 3) float32 shared vertices
    vnorm: Optional[np.ndarray] = None  # (m, 3) float32 shared normals
    faces: Optional[np.ndarray] = None  # (n, 3) int vertex indices
    uv: Optional[np.ndarray] = None     # (n, 3, 2) float32 texture coordinates (soup form)


@dataclass
class SceneSpec:
    camera: dict
    materials: List[dict]
    meshes: List[MeshSpec]
    lights: List[dict]
    textures: List[dict] = field(default_factory=list)
    texture_images: dict = field(default_factory=dict)  # texture file name -> (h, w, 4) uint8, row 0 = top


def _quad(p0, p1, p2, p3, normal) -> np.ndarray:
    """Two triangles of a planar quad, wound so (v1-v0)x(v2-v0) points along `normal`."""
    p = np.array([p0, p1, p2, p3], np.float64)
    tris = [[p[0], p[1], p[2]], [p[0], p[2], p[3]]]
    out = []
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        if np.dot(n, normal) < 0:
            t = [t[0], t[2], t[1]]
        out.append(t)
    return np.array(out, np.float32)


def _box(lo, hi) -> np.ndarray:
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        _quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
        _quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1), (0, 1, 0)),
        _quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1), (-1, 0, 0)),
        _quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0)),
        _quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (0, 0, -1)),
        _quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1)),
    ]
    return np.concatenate(faces)


def cornell_shell() -> List[MeshSpec]:
    """Open-front Cornell box x in [-1,1], y in [0,2], z in [-1,1]; the camera
    (orbit direction (0,0,-1)) looks along +z, so image-left is +x."""
    floor = _quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1), (0, 1, 0))
    ceil = _quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0))
    back = _quad((-1, 0, 1), (1, 0, 1), (1, 2, 1), (-1, 2, 1), (0, 0, -1))
    left = _quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0))     # red, image left
    right = _quad((-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1), (1, 0, 0))  # green, image right
    light = _quad((-0.25, 1.98, -0.2), (0.25, 1.98, -0.2), (0.25, 1.98, 0.3), (-0.25, 1.98, 0.3), (0, -1, 0))
    short_box = _box((-0.75, 0.0, -0.6), (-0.15, 0.6, 0.0))
    tall_box = _box((0.1, 0.0, 0.1), (0.7, 1.2, 0.7))
    return [
        MeshSpec("floor", "white", floor),
        MeshSpec("ceiling", "white", ceil),
        MeshSpec("back", "white", back),
        MeshSpec("left", "red", left),
        MeshSpec("right", "green", right),
        MeshSpec("short_box", "white", short_box),
        MeshSpec("tall_box", "white", tall_box),
        MeshSpec("light", "light", light),
    ]


def _base_materials() -> List[dict]:
    return [
        {"name": "white", "type": "diffuse", "basecolor": list(WHITE)},
        {"name": "red", "type": "diffuse", "basecolor": list(RED)},
        {"name": "green", "type": "diffuse", "basecolor": list(GREEN)},
        {"name": "light", "type": "light", "emission": list(LIGHT_RADIANCE)},
    ]


def _camera(w, h, spp, depth, rr=0.8, rr_start=3, seed=1234) -> dict:
    return {
        "type": "pinhole",
        "resolution": [w, h],
        "fov": 39.3077,
        "max_bounces": depth,
        "rr": rr,
        "spp": spp,
        "direction": [0, 0, -1],
        "denoise": False,
        "seed": seed,
        "rr_start_depth": rr_start,
    }


def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: (vertices (m,3) float64, faces (n,3) int) with outward winding."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces, np.int64)
    # Outward winding: (v1-v0)x(v2-v0) along the centroid direction.
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    c = V[F].mean(axis=1)
    flip = (n * c).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def cornell_small(seed: int = 1) -> SceneSpec:
    return SceneSpec(camera=_camera(256, 256, 16, 8), materials=_base_materials(), meshes=cornell_shell(),
                     lights=[{"name": "light0", "type": "area", "lightMeshName": "light", "material": "light"}])


def cornell_spheres(seed: int = 2, width: int = 512, height: int = 512, spp: int = 64, depth: int = 12) -> SceneSpec:
    rng = np.random.default_rng(seed)
    V, F = icosphere(4)
    meshes = cornell_shell()
    mats = _base_materials()
    boxes = [((-0.75, -0.6), (-0.15, 0.0)), ((0.1, 0.1), (0.7, 0.7))]
    placed = []
    kinds = ["pbr", "pbr", "pbr", "dielectric"]
    for k, kind in enumerate(kinds):
        r = float(rng.uniform(0.15, 0.25))
        for _ in range(10000):
            cx, cz = rng.uniform(-0.95 + r, 0.95 - r, size=2)
            ok = all(not (lo[0] - r < cx < hi[0] + r and lo[1] - r < cz < hi[1] + r) for lo, hi in boxes)
            ok = ok and all((cx - px) ** 2 + (cz - pz) ** 2 > (r + pr + 0.02) ** 2 for px, pz, pr in placed)
            if ok:
                break
        else:
            raise RuntimeError("could not place sphere")
        placed.append((cx, cz, r))
        center = np.array([cx, r + 1e-3, cz])
        pos = (V[F] * r + center).astype(np.float32)
        nrm = V[F].astype(np.float32)
        name = f"sphere{k}"
        if kind == "pbr":
            metallic = float(rng.integers(0, 2))
            rough = float(rng.uniform(0.05, 0.8))
            base = [float(x) for x in rng.uniform(0.2, 0.95, size=3)]
            mats.append({"name": f"mat_{name}", "type": "pbr", "basecolor": base, "metallic": metallic,
                         "roughness": rough})
        else:
            mats.append({"name": f"mat_{name}", "type": "dielectric", "basecolor": [1.0, 1.0, 1.0], "ior": 1.5})
        meshes.append(MeshSpec(name, f"mat_{name}", pos, nrm))
    return SceneSpec(camera=_camera(width, height, spp, depth), materials=mats, meshes=meshes,
                     lights=[{"name": "light0", "type": "area", "lightMeshName": "light", "material": "light"}])


def _fmt(x: float) -> str:
    return "%.9g" % float(x)


def write_obj(path: str, mesh: MeshSpec) -> None:
    if mesh.faces is not None:
        _write_indexed_obj(path, mesh)
        return
    n = mesh.pos.shape[0]
    lines = [f"o {mesh.name}"]
    for t in range(n):
        for k in range(3):
            p = mesh.pos[t, k]
            lines.append(f"v {_fmt(p[0])} {_fmt(p[1])} {_fmt(p[2])}")
    if mesh.uv is not None:
        for t in range(n):
            for k in range(3):
                q = mesh.uv[t, k]
                lines.append(f"vt {_fmt(q[0])} {_fmt(q[1])}")
    if mesh.nrm is not None:
        for t in range(n):
            for k in range(3):
                q = mesh.nrm[t, k]
                lines.append(f"vn {_fmt(q[0])} {_fmt(q[1])} {_fmt(q[2])}")
        for t in range(n):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            if mesh.uv is not None:
                lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}")
            else:
                lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
    elif mesh.uv is not None:
        for t in range(n):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            lines.append(f"f {a}/{a} {b}/{b} {c}/{c}")
    else:
        for t in range(n):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            lines.append(f"f {a} {b} {c}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _write_indexed_obj(path: str, mesh: MeshSpec) -> None:
    # float32 values go through Python floats (exact) and %.9g (round-trips float32).
    with open(path, "w") as f:
        f.write(f"o {mesh.name}\n")
        f.write("\n".join(["v %.9g %.9g %.9g" % tuple(r) for r in mesh.verts.astype(np.float32).tolist()]) + "\n")
        fi = (mesh.faces.astype(np.int64) + 1).tolist()
        if mesh.vnorm is not None:
            f.write("\n".join(["vn %.9g %.9g %.9g" % tuple(r) for r in mesh.vnorm.astype(np.float32).tolist()])
                    + "\n")
            f.write("\n".join(["f %d//%d %d//%d %d//%d" % (a, a, b, b, c, c) for a, b, c in fi]) + "\n")
        else:
            f.write("\n".join(["f %d %d %d" % (a, b, c) for a, b, c in fi]) + "\n")


def write_png_rgba(path: str, rgba: np.ndarray) -> None:
    """8-bit RGBA PNG, row 0 = top of the image (no row filters, one data chunk)."""
    import struct
    import zlib

    h, w, _ = rgba.shape
    raw = b"".join(b"\x00" + np.ascontiguousarray(rgba[y], np.uint8).tobytes() for y in range(h))

    def chunk(tag: bytes, body: bytes) -> bytes:
        return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
                + chunk(b"IDAT", zlib.compress(raw, 9)) + chunk(b"IEND", b""))


def _planar_uv(pos: np.ndarray, lo: float, hi: float) -> np.ndarray:
    """Texture coordinates of a planar triangle soup (n, 3, 3): the two axes
    across the dominant normal axis, each mapped from the mesh's extent to [lo, hi]."""
    p = pos.reshape(-1, 3).astype(np.float64)
    n = np.abs(np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0]).sum(axis=0))
    axes = [a for a in range(3) if a != int(np.argmax(n))]
    uv = np.empty((p.shape[0], 2))
    for k, a in enumerate(axes):
        span = max(p[:, a].max() - p[:, a].min(), 1e-9)
        uv[:, k] = lo + (hi - lo) * (p[:, a] - p[:, a].min()) / span
    return uv.reshape(pos.shape[0], 3, 2).astype(np.float32)


def _sphere_uv(dirs: np.ndarray) -> np.ndarray:
    """Longitude / latitude coordinates of unit directions (n, 3, 3)."""
    u = 0.5 + np.arctan2(dirs[..., 2], dirs[..., 0]) / (2 * np.pi)
    v = 0.5 + np.arcsin(np.clip(dirs[..., 1], -1, 1)) / np.pi
    return np.stack([u, v], axis=-1).astype(np.float32)


def _texture_image(h: int, w: int, rng) -> np.ndarray:
    """Checker pattern with noise: every texel differs from its neighbours."""
    y, x = np.mgrid[0:h, 0:w]
    check = ((x // 3 + y // 2) % 2)[..., None] * np.array([150, 60, 20])
    noise = rng.integers(0, 100, (h, w, 3))
    rgb = np.clip(40 + check + noise, 0, 255)
    return np.concatenate([rgb, rng.integers(0, 256, (h, w, 1))], axis=2).astype(np.uint8)


def cornell_sss_walk(width: int = 96, height: int = 96, spp: int = 4, depth: int = 8) -> SceneSpec:
    """The Cornell shell with two random-walk subsurface spheres of
    chromatic media (optical radii from about 1.5 to 8), one of them resting
    against the short box so that walks also leave through faces of other
    meshes inside the sphere."""
    V, F = icosphere(3)
    meshes = cornell_shell()
    mats = _base_materials()
    mats.append({"name": "jade", "type": "subsurface", "basecolor": [1, 1, 1], "ior": 1.45, "roughness": 0.2,
                 "subsurface_method": "random_walk", "sigma_a": [2.0, 0.3, 1.5], "sigma_s": [6.0, 14.0, 9.0]})
    mats.append({"name": "milk", "type": "subsurface", "basecolor": [1, 1, 1], "ior": 1.3, "roughness": 0.0,
                 "subsurface_method": "random_walk", "sigma_a": [0.05, 0.1, 0.4], "sigma_s": [20.0, 22.0, 26.0]})
    r0, r1 = 0.3, 0.22
    meshes.append(MeshSpec("jade_ball", "jade", (V[F] * r0 + np.array([0.45, r0 + 1e-3, -0.45])).astype(np.float32),
                           V[F].astype(np.float32)))
    meshes.append(MeshSpec("milk_ball", "milk", (V[F] * r1 + np.array([-0.45, 0.6 + r1 * 0.8, -0.3])).astype(np.float32),
                           V[F].astype(np.float32)))
    return SceneSpec(camera=_camera(width, height, spp, depth), materials=mats, meshes=meshes,
                     lights=[{"name": "light0", "type": "area", "lightMeshName": "light", "material": "light"}])


def textured_cornell(seed: int = 5, width: int = 96, height: int = 96, spp: int = 4, depth: int = 6) -> SceneSpec:
    """The Cornell shell with diffuse textures on every lobe that reads an albedo:
    a diffuse back wall whose coordinates run past [0, 1] (clamped), a GGX floor,
    a metallic GGX tall box, a 1x1-textured ceiling, a subsurface sphere and a
    textured dielectric sphere (whose texture the reference ignores).  Sphere
    longitudes wrap across u = 0."""
    rng = np.random.default_rng(seed)
    meshes = cornell_shell()
    byname = {m.name: m for m in meshes}
    images = {"wall.png": _texture_image(23, 37, rng), "floor.png": _texture_image(64, 64, rng),
              "tiny.png": _texture_image(3, 5, rng), "one.png": _texture_image(1, 1, rng),
              "sphere.png": _texture_image(16, 32, rng)}
    textures = [{"name": n.split(".")[0], "path": n} for n in images]
    mats = _base_materials() + [
        {"name": "wall_tex", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5], "diffuse_texture": "wall"},
        {"name": "floor_tex", "type": "pbr", "basecolor": [0.5, 0.5, 0.5], "metallic": 0.0, "roughness": 0.45,
         "diffuse_texture": "floor"},
        {"name": "box_tex", "type": "pbr", "basecolor": [0.5, 0.5, 0.5], "metallic": 1.0, "roughness": 0.3,
         "diffuse_texture": "tiny"},
        {"name": "short_tex", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5], "diffuse_texture": "tiny"},
        {"name": "ceiling_tex", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5], "diffuse_texture": "one"},
        {"name": "sss_tex", "type": "subsurface", "basecolor": [0.8, 0.8, 0.8], "scatter_distance": [0.3, 0.15, 0.08],
         "ior": 1.4, "roughness": 0.3, "diffuse_texture": "sphere"},
        {"name": "glass_tex", "type": "dielectric", "basecolor": [0.9, 0.95, 1.0], "ior": 1.5,
         "diffuse_texture": "sphere"},
    ]
    for name, mat, lo, hi in (("back", "wall_tex", -0.3, 1.25), ("floor", "floor_tex", 0.0, 1.0),
                              ("ceiling", "ceiling_tex", 0.0, 1.0)):
        byname[name].material = mat
        byname[name].uv = _planar_uv(byname[name].pos, lo, hi)
    for name, mat in (("tall_box", "box_tex"), ("short_box", "short_tex")):
        m = byname[name]
        m.material = mat
        m.uv = np.concatenate([_planar_uv(m.pos[2 * f:2 * f + 2], 0.0, 1.0) for f in range(m.pos.shape[0] // 2)])
    V, F = icosphere(3)
    for name, mat, center, r in (("sss_sphere", "sss_tex", (-0.45, 0.95, -0.3), 0.22),
                                 ("glass_sphere", "glass_tex", (0.45, 1.45, 0.35), 0.2)):
        dirs = V[F]
        pos = (dirs * r + np.array(center)).astype(np.float32)
        meshes.append(MeshSpec(name, mat, pos, dirs.astype(np.float32), uv=_sphere_uv(dirs)))
    return SceneSpec(camera=_camera(width, height, spp, depth), materials=mats, meshes=meshes,
                     lights=[{"name": "light0", "type": "area", "lightMeshName": "light", "material": "light"}],
                     textures=textures, texture_images=images)


def write_pfm(path: str, rgb: np.ndarray) -> None:
    """Little-endian colour PFM; rgb is (h, w, 3) with row 0 = top of the image."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(rgb[::-1].astype("<f4")).tobytes())


def _rle_channel(vals: np.ndarray) -> bytes:
    """One channel of a new-style RGBE scanline: runs of 3..127 equal bytes as
    (128 + count, value), everything else as literal dumps of at most 128."""
    out = bytearray()
    n = len(vals)
    # start of each stretch of equal values and its length
    edges = np.flatnonzero(np.diff(vals.astype(np.int16))) + 1
    starts = np.concatenate([[0], edges])
    lens = np.diff(np.concatenate([starts, [n]]))
    lit = bytearray()

    def flush():
        for k in range(0, len(lit), 128):
            chunk = lit[k:k + 128]
            out.append(len(chunk))
            out.extend(chunk)
        lit.clear()

    for st, ln in zip(starts.tolist(), lens.tolist()):
        v = int(vals[st])
        while ln >= 3:
            run = min(ln, 127)
            flush()
            out += bytes((128 + run, v))
            ln -= run
        lit.extend(bytes([v]) * ln)
    flush()
    return bytes(out)


def write_hdr(path: str, rgb: np.ndarray) -> None:
    """Radiance RGBE file (the reference's environment-map format, read with
    stbi_loadf): Ward's shared-exponent encoding of (h, w, 3) float RGB, row 0 =
    first scanline, new-style run-length scanlines (widths 8..32767) or flat
    pixels otherwise."""
    h, w, _ = rgb.shape
    v = rgb.max(axis=2).astype(np.float64)
    m, e = np.frexp(v)
    ok = v > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    px = np.zeros((h, w, 4), np.uint8)
    px[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    px[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
        if w < 8 or w >= 32768:
            f.write(px.tobytes())
            return
        for y in range(h):
            f.write(bytes((2, 2, w >> 8, w & 255)) + b"".join(_rle_channel(px[y, :, c]) for c in range(4)))


def sky_env(width: int = 2048, height: int = 1024, sun_dir=(0.35, 0.6, -0.72), sun_peak: float = 1e4,
            sun_radius_deg: float = 0.5) -> np.ndarray:
    """Procedural HDR sky: zenith-to-horizon gradient, dark ground, Gaussian sun
    lobe of the given peak radiance whose 1-sigma radius is `sun_radius_deg`.
    Pixel (x, y) looks along theta = (y + 0.5) / h * pi from +y, phi = (x + 0.5) / w * 2 pi
    with direction (sin theta sin phi, cos theta, sin theta cos phi)."""
    y = (np.arange(height) + 0.5) / height * np.pi
    x = (np.arange(width) + 0.5) / width * 2 * np.pi
    th, ph = np.meshgrid(y, x, indexing="ij")
    d = np.stack([np.sin(th) * np.sin(ph), np.cos(th), np.sin(th) * np.cos(ph)], axis=-1)
    up = np.clip(d[..., 1], 0, 1)[..., None]
    zenith = np.array([0.25, 0.45, 1.0])
    horizon = np.array([1.0, 0.95, 0.9])
    sky = horizon * (1 - up) ** 3 + zenith * (1 - (1 - up) ** 3)
    ground = np.array([0.12, 0.1, 0.08])
    rgb = np.where(d[..., 1:2] >= 0, sky, ground)
    s = np.array(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    ang = np.arccos(np.clip((d * s).sum(-1), -1, 1))
    lobe = sun_peak * np.exp(-0.5 * (ang / np.radians(sun_radius_deg)) ** 2)
    rgb = rgb + lobe[..., None] * np.array([1.0, 0.95, 0.85])
    return rgb.astype(np.float32)


def _fbm(n: int, seed: int, octaves: int = 6) -> np.ndarray:
    """Fractional Brownian motion heights on an n x n grid (value noise, smoothstep)."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n)
    h = np.zeros((n, n))
    amp, freq = 1.0, 4
    for _ in range(octaves):
        lat = rng.standard_normal((freq + 2, freq + 2))
        g = u * freq
        i0 = np.floor(g).astype(int)
        t = g - i0
        t = t * t * (3 - 2 * t)
        a = lat[i0][:, i0]
        b = lat[i0 + 1][:, i0]
        c = lat[i0][:, i0 + 1]
        d = lat[i0 + 1][:, i0 + 1]
        tx, ty = t[None, :], t[:, None]
        h += amp * ((a * (1 - ty) + b * ty) * (1 - tx) + (c * (1 - ty) + d * ty) * tx)
        amp *= 0.5
        freq *= 2
    return h


def terrain_mesh(grid: int, size: float, height: float, seed: int) -> MeshSpec:
    """(grid-1)^2 * 2 triangles over [-size/2, size/2]^2, smooth normals, faces up."""
    h = _fbm(grid, seed) * height
    u = np.linspace(-size / 2, size / 2, grid)
    X, Z = np.meshgrid(u, u)  # row index = z
    verts = np.stack([X, h, Z], axis=-1).reshape(-1, 3)
    gz, gx = np.gradient(h, u, u)
    nrm = np.stack([-gx, np.ones_like(h), -gz], axis=-1).reshape(-1, 3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    idx = np.arange(grid * grid).reshape(grid, grid)
    a, b, c, d = idx[:-1, :-1], idx[:-1, 1:], idx[1:, :-1], idx[1:, 1:]
    # (v1-v0) x (v2-v0) points +y for (a, c, b) and (b, c, d) with x along columns, z along rows.
    faces = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([b, c, d], -1).reshape(-1, 3)])
    return MeshSpec("terrain", "terrain", verts=verts.astype(np.float32), vnorm=nrm.astype(np.float32), faces=faces)


def burley_from_coefficients(sigma_a, sigma_s):
    """Map volume coefficients (1/m) to the reference's Burley parameters:
    basecolor = single-scattering albedo, scatter_distance = mean free path / s(A)
    with Christensen & Burley's (2015) fit s(A) = 1.85 - A + 7 |A - 0.8|^3."""
    sa, ss = np.asarray(sigma_a, float), np.asarray(sigma_s, float)
    albedo = ss / (sa + ss)
    s = 1.85 - albedo + 7 * np.abs(albedo - 0.8) ** 3
    return albedo, (1.0 / (sa + ss)) / s


def terrain_env(seed: int = 3, grid: int = 1001, size: float = 100.0, env_w: int = 2048, env_h: int = 1024,
                width: int = 1920, height: int = 1080, spp: int = 32, depth: int = 16,
                sphere_subdiv: int = 4, env_file: str = "sky.pfm", sss: str = "random_walk") -> SceneSpec:
    """Terrain + subsurface sphere under a procedural HDR sky.  env_file picks
    the map's format by extension: .pfm keeps FP32 texels, .hdr is the
    reference's Radiance RGBE.  The sphere's medium has sigma_a, sigma_s drawn
    log-uniform in [0.1, 10] per metre (scene unit) per channel; sss picks its
    transport: "random_walk" through those coefficients, or "burley", the
    reference's diffusion profile with parameters mapped from them
    (burley_from_coefficients)."""
    rng = np.random.default_rng(seed)
    terrain = terrain_mesh(grid, size, height=size * 0.06, seed=seed)
    V, F = icosphere(sphere_subdiv)
    r = size * 0.04
    # Rest the sphere on the terrain near the centre.
    c = grid // 2
    y0 = float(terrain.verts.reshape(grid, grid, 3)[c, c, 1])
    sphere = MeshSpec("sss_ball", "sss", pos=(V[F] * r + np.array([0.0, y0 + r * 0.9, 0.0])).astype(np.float32),
                      nrm=V[F].astype(np.float32))
    sig_a = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 3))
    sig_s = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 3))
    albedo, dist = burley_from_coefficients(sig_a, sig_s)
    if sss not in ("random_walk", "burley"):
        raise ValueError(f"unknown subsurface method {sss!r}")
    mats = [
        {"name": "terrain", "type": "diffuse", "basecolor": [0.35, 0.3, 0.22]},
        {"name": "sss", "type": "subsurface", "basecolor": [float(x) for x in albedo],
         "scatter_distance": [float(x) for x in dist], "ior": 1.4, "roughness": 0.3,
         "subsurface_method": sss, "sigma_a": [float(x) for x in sig_a], "sigma_s": [float(x) for x in sig_s]},
    ]
    spec = SceneSpec(camera=dict(_camera(width, height, spp, depth), direction=[0.35, 0.45, -1.0]),
                     materials=mats, meshes=[terrain, sphere],
                     lights=[{"name": "sky", "type": "env", "path": env_file}])
    spec.env = sky_env(env_w, env_h)
    return spec


def batch_scene(seed: int, target_tris: int, width: int, spp: int, depth: int = 12) -> SceneSpec:
    """One scene of the uneven batch: the Cornell shell, up to four icospheres
    (subdivision 4 or 3) with the materials of the spheres scene (GGX
    metallic-roughness, IOR-1.5 dielectric, Lambertian), and the remaining
    triangle budget as a heightfield patch on the floor."""
    rng = np.random.default_rng(seed)
    meshes = cornell_shell()
    mats = _base_materials()
    budget = max(0, int(target_tris) - 36)
    boxes = [((-0.75, -0.6), (-0.15, 0.0)), ((0.1, 0.1), (0.7, 0.7))]
    placed = []
    kinds = ["pbr", "dielectric", "diffuse", "pbr"]
    for k, kind in enumerate(kinds):
        sub = 4 if budget >= 2 * 5120 else (3 if budget >= 2 * 1280 else None)
        if sub is None:
            break
        V, F = icosphere(sub)
        r = float(rng.uniform(0.12, 0.22))
        for _ in range(1000):
            cx, cz = rng.uniform(-0.95 + r, 0.95 - r, size=2)
            ok = all(not (lo[0] - r < cx < hi[0] + r and lo[1] - r < cz < hi[1] + r) for lo, hi in boxes)
            ok = ok and all((cx - px) ** 2 + (cz - pz) ** 2 > (r + pr + 0.02) ** 2 for px, pz, pr in placed)
            if ok:
                break
        else:
            break
        placed.append((cx, cz, r))
        name = f"sphere{k}"
        if kind == "pbr":
            mats.append({"name": f"mat_{name}", "type": "pbr",
                         "basecolor": [float(x) for x in rng.uniform(0.2, 0.95, size=3)],
                         "metallic": float(rng.integers(0, 2)), "roughness": float(rng.uniform(0.05, 0.8))})
        elif kind == "dielectric":
            mats.append({"name": f"mat_{name}", "type": "dielectric", "basecolor": [1.0, 1.0, 1.0], "ior": 1.5})
        else:
            mats.append({"name": f"mat_{name}", "type": "diffuse",
                         "basecolor": [float(x) for x in rng.uniform(0.2, 0.9, size=3)]})
        center = np.array([cx, r + 0.06, cz])
        meshes.append(MeshSpec(name, f"mat_{name}", (V[F] * r + center).astype(np.float32), V[F].astype(np.float32)))
        budget -= len(F)
    if budget >= 2:
        g = int(np.sqrt(budget / 2.0)) + 1
        patch = terrain_mesh(g, 1.9, 0.05, seed)
        v = patch.verts.copy()
        v[:, 1] -= v[:, 1].min() - 0.002
        mats.append({"name": "patch", "type": "diffuse", "basecolor": [float(x) for x in rng.uniform(0.3, 0.8, 3)]})
        meshes.append(MeshSpec("patch", "patch", verts=v, vnorm=patch.vnorm, faces=patch.faces))
    return SceneSpec(camera=_camera(width, width, spp, depth), materials=mats, meshes=meshes,
                     lights=[{"name": "light0", "type": "area", "lightMeshName": "light", "material": "light"}])


def batch_mix(n: int = 64, seed: int = 4, lo_tris: float = 1e3, hi_tris: float = 5e5):
    """The uneven batch: n seeded scenes with triangle counts log-uniform in
    [lo_tris, hi_tris], square resolutions from {128, 256, 512, 1024} and spp
    uniform on 1..64.  Returns [(name, SceneSpec)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        tris = float(np.exp(rng.uniform(np.log(lo_tris), np.log(hi_tris))))
        width = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.integers(1, 65))
        out.append((f"batch{k:02d}", batch_scene(1000 + k, int(tris), width, spp)))
    return out


def write_scene(spec: SceneSpec, out_dir: str, name: str = "scene") -> str:
    """Write <out_dir>/<name>.json and one OBJ per mesh; returns the JSON path."""
    os.makedirs(out_dir, exist_ok=True)
    objects = []
    for m in spec.meshes:
        fn = f"{name}_{m.name}.obj"
        write_obj(os.path.join(out_dir, fn), m)
        objects.append({"name": m.name, "path": fn, "translate": [0, 0, 0], "scale": [1, 1, 1],
                        "rotate": [{"rotateX": 0}], "material": m.material})
    for fn, img in spec.texture_images.items():
        write_png_rgba(os.path.join(out_dir, fn), img)
    env = getattr(spec, "env", None)
    if env is not None:
        for L in spec.lights:
            if L.get("type") == "env":
                writer = write_hdr if L["path"].lower().endswith(".hdr") else write_pfm
                writer(os.path.join(out_dir, L["path"]), env)
    doc = {"camera": spec.camera, "textures": spec.textures, "materials": spec.materials, "objects": objects,
           "lights": spec.lights}
    path = os.path.join(out_dir, f"{name}.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path


WORKLOADS = {
    "cornell_small": cornell_small,
    "cornell_spheres": cornell_spheres,
    "textured_cornell": textured_cornell,
    "cornell_sss_walk": cornell_sss_walk,
}
# Large scenes are generated on demand (the full terrain is 2M triangles).
LARGE_WORKLOADS = {
    "terrain_env": terrain_env,
}
