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
    pos: np.ndarray                    # (n, 3, 3) float32
    nrm: Optional[np.ndarray] = None   # (n, 3, 3) float32 per-vertex normals or None


@dataclass
class SceneSpec:
    camera: dict
    materials: List[dict]
    meshes: List[MeshSpec]
    lights: List[dict]
    textures: List[dict] = field(default_factory=list)


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
    n = mesh.pos.shape[0]
    lines = [f"o {mesh.name}"]
    for t in range(n):
        for k in range(3):
            p = mesh.pos[t, k]
            lines.append(f"v {_fmt(p[0])} {_fmt(p[1])} {_fmt(p[2])}")
    if mesh.nrm is not None:
        for t in range(n):
            for k in range(3):
                q = mesh.nrm[t, k]
                lines.append(f"vn {_fmt(q[0])} {_fmt(q[1])} {_fmt(q[2])}")
        for t in range(n):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            lines.append(f"f {a}//{a} {b}//{b} {c}//{c}")
    else:
        for t in range(n):
            a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
            lines.append(f"f {a} {b} {c}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_scene(spec: SceneSpec, out_dir: str, name: str = "scene") -> str:
    """Write <out_dir>/<name>.json and one OBJ per mesh; returns the JSON path."""
    os.makedirs(out_dir, exist_ok=True)
    objects = []
    for m in spec.meshes:
        fn = f"{name}_{m.name}.obj"
        write_obj(os.path.join(out_dir, fn), m)
        objects.append({"name": m.name, "path": fn, "translate": [0, 0, 0], "scale": [1, 1, 1],
                        "rotate": [{"rotateX": 0}], "material": m.material})
    doc = {"camera": spec.camera, "textures": spec.textures, "materials": spec.materials, "objects": objects,
           "lights": spec.lights}
    path = os.path.join(out_dir, f"{name}.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path


WORKLOADS = {
    "cornell_small": cornell_small,
    "cornell_spheres": cornell_spheres,
}
