"""OBJ import of the library (parsed in chunks on host threads) against the
oracle's sequential parser (OBJLoader::LoadObject semantics,
Src/Utils/objloader.cpp:66-210): polygons fanned into triangles, negative
(relative) indices, faces with and without texture / normal indices,
object / group / material statements regrouping faces into meshes, CRLF line
ends, comments, and a file spanning several parse chunks, with relative
indices crossing the chunk boundaries."""
import json
import os

import numpy as np

import cudaraytracer_amd as rt
from oracle import rtref


def _random_obj(seed, nv, nf):
    r = np.random.default_rng(seed)
    lines = ["# generated", "mtllib none.mtl"]
    lines += ["v %.6f %.6f %.6f" % tuple(x) for x in r.uniform(-1, 1, (nv, 3))]
    lines += ["vn %.5f %.5f %.5f" % tuple(x) for x in r.normal(size=(nv // 2 + 1, 3))]
    lines += ["vt %.4f %.4f" % tuple(x) for x in r.uniform(0, 1, (nv // 3 + 1, 2))]
    nvt, nvn = nv // 3 + 1, nv // 2 + 1
    for _ in range(nf):
        u = r.random()
        if u < 0.01:
            lines.append(str(r.choice(["o", "g"])) + " " + str(r.choice(["a", "b", "part two", "a"])))
        elif u < 0.02:
            lines.append("usemtl " + str(r.choice(["red", "m", "other"])))
        elif u < 0.025:
            lines.append("   # comment")
        elif u < 0.03:
            lines.append("v %.3f %.3f %.3f" % tuple(r.uniform(0, 1, 3)))
            nv += 1
        toks = []
        for _ in range(int(r.choice([3, 3, 3, 4, 5]))):
            vi = int(r.integers(1, nv + 1))
            v = str(vi) if r.random() < 0.8 else str(vi - nv - 1)  # negative = relative
            ti, ni = int(r.integers(1, nvt + 1)), int(r.integers(1, nvn + 1))
            toks.append([v, f"{v}/{ti}", f"{v}//{ni}", f"{v}/{ti}/{ni}"][int(r.integers(0, 4))])
        lines.append("f " + " ".join(toks))
    return "\r\n".join(lines) + "\r\n"


def test_obj_import_matches_oracle_parser(tmp_path):
    text = _random_obj(3, 20000, 40000)
    assert len(text) > (3 << 20) // 2  # more than one 1 MiB parse chunk
    (tmp_path / "m.obj").write_bytes(text.encode())
    doc = {
        "camera": {"type": "pinhole", "resolution": [8, 8], "fov": 40, "spp": 1, "max_bounces": 2, "rr": 0.8,
                   "direction": [0, 0, -1]},
        "textures": [],
        "materials": [{"name": "m", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5]},
                      {"name": "red", "type": "diffuse", "basecolor": [0.9, 0.1, 0.1]}],
        "objects": [{"name": "obj", "path": "m.obj", "default_material": "m",
                     "translate": [0.1, -0.2, 0.3], "rotate": [{"rotateY": 30}], "scale": [1.5, 1, 0.5]}],
        "lights": [],
    }
    path = str(tmp_path / "s.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    scene = rt.Scene(path, bvh_builder=rt.BVH_HOST)
    pos, nrm, uv, mesh = scene.triangles()
    a = rtref.parse_scene(path)
    assert pos.shape[0] == a.pos.shape[0] > 40000
    assert np.array_equal(pos.view(np.int32), a.pos.view(np.int32))
    assert np.array_equal(nrm.view(np.int32), a.nrm.view(np.int32))
    assert np.array_equal(uv.view(np.int32), a.uv.view(np.int32))
    assert np.array_equal(mesh, a.mesh)
    assert scene.info().num_meshes == len(a.mesh_material) > 3
