"""OBJ import and object transforms against the reference's own libraries.

The reference reads OBJ files with its vendored tinyobjloader and transforms
them with Eigen 3.4.0 (Src/Utils/objloader.cpp:66-210, Src/Core/scene.cpp:
405-448).  oracle/objload_ref.cpp drives those two libraries, built from the
reference tree (`make -C oracle ref`; skipped where the tree is absent), the
way the reference does; the library's importer and the oracle's restatement
(oracle/rtref_obj.py) must return the same triangles bit for bit on generated
files that exercise tinyobj's number reader, its ear clipping, its shape and
material rules, relative indices, line endings and the T * R * S transform
with its inverse-transpose normals."""
import ctypes
import json
import os

import numpy as np
import pytest

import cudaraytracer_amd as rt
from oracle import rtref_obj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libobjref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference tinyobj/Eigen build not present")

_L = None


def ref():
    global _L
    if _L is None:
        L = ctypes.CDLL(REF_LIB)
        vp = ctypes.c_void_p
        L.ref_obj_load.restype = vp
        L.ref_obj_load.argtypes = [ctypes.c_char_p, vp, vp, vp, ctypes.c_int, vp]
        for f in ("ref_obj_num_triangles", "ref_obj_num_meshes"):
            getattr(L, f).argtypes = [vp]
        L.ref_obj_triangles.argtypes = [vp, vp, vp, vp, vp]
        L.ref_obj_mesh_name.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.ref_obj_free.argtypes = [vp]
        L.ref_object_transform.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp, vp]
        L.ref_inverse4.argtypes = [vp, vp]
        _L = L
    return _L


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def ref_load(path, t, rots, s):
    axes = np.array([a for a, _ in rots] or [0], np.int32)
    degs = np.array([d for _, d in rots] or [0], np.float32)
    t = np.array(t, np.float32)
    s = np.array(s, np.float32)
    h = ref().ref_obj_load(path.encode(), t.ctypes.data, axes.ctypes.data, degs.ctypes.data, len(rots), s.ctypes.data)
    if not h:
        return None
    try:
        n = ref().ref_obj_num_triangles(h)
        pos, nrm = np.zeros((n, 9), np.float32), np.zeros((n, 9), np.float32)
        uv, mesh = np.zeros((n, 6), np.float32), np.zeros(n, np.int32)
        ref().ref_obj_triangles(h, pos.ctypes.data, nrm.ctypes.data, uv.ctypes.data, mesh.ctypes.data)
        names = []
        for m in range(ref().ref_obj_num_meshes(h)):
            a, b = ctypes.create_string_buffer(4096), ctypes.create_string_buffer(4096)
            ref().ref_obj_mesh_name(h, m, a, b, 4096)
            names.append((a.value.decode("latin-1"), b.value.decode("latin-1")))
        return pos, nrm, uv, mesh, names
    finally:
        ref().ref_obj_free(h)


def _obj_dict(t, rots, s):
    return {"translate": [float(x) for x in t], "scale": [float(x) for x in s],
            "rotate": [{["rotateX", "rotateY", "rotateZ"][a]: float(v)} for a, v in rots]}


def oracle_load(path, t, rots, s):
    xf, inv = rtref_obj.object_transform(_obj_dict(t, rots, s))
    meshes = rtref_obj.load_obj(path, xf, inv)
    if not meshes:
        return None
    pos = np.concatenate([m["pos"] for m in meshes])
    nrm = np.concatenate([m["nrm"] for m in meshes])
    uv = np.concatenate([m["uv"] for m in meshes])
    mesh = np.concatenate([np.full(len(m["pos"]), k, np.int32) for k, m in enumerate(meshes)])
    return pos, nrm, uv, mesh, [(m["name"], m["material"]) for m in meshes]


def library_load(tmp, obj_path, t, rots, s, mat_names):
    """The library's loader on a one-object scene whose meshes bind through
    the OBJ's material names, or a default material."""
    mats = [{"name": "dm", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5]}]
    mats += [{"name": n, "type": "diffuse", "basecolor": [0.2, 0.3, 0.4]} for n in mat_names if n and n != "dm"]
    doc = {"camera": {"type": "pinhole", "resolution": [8, 8], "fov": 40, "spp": 1, "max_bounces": 2, "rr": 0.8},
           "textures": [], "materials": mats,
           "objects": [dict(_obj_dict(t, rots, s), name="o", path=obj_path, default_material="dm")], "lights": []}
    jp = os.path.join(tmp, "scene.json")
    with open(jp, "w") as f:
        json.dump(doc, f)
    try:
        scene = rt.Scene(jp, bvh_builder=rt.BVH_HOST)
    except rt.RtError as e:
        if "no triangles" in str(e):
            return None
        raise
    info = scene.info()
    n = info.num_triangles
    pos, nrm, uv = (np.zeros((n, k), np.float32) for k in (9, 9, 6))
    fp = ctypes.POINTER(ctypes.c_float)
    rt.lib().rt_scene_get_triangles(scene._h, pos.ctypes.data_as(fp), nrm.ctypes.data_as(fp), uv.ctypes.data_as(fp),
                                    None)
    mesh4, _ = scene.meshes()
    index = {m["name"]: i for i, m in enumerate(mats)}
    return pos, nrm, uv, mesh4, index


def check_all(tmp, obj_path, t=(0, 0, 0), rots=(), s=(1, 1, 1), expect_tris=True):
    rots = list(rots)
    r = ref_load(obj_path, t, rots, s)
    o = oracle_load(obj_path, t, rots, s)
    if r is None or len(r[0]) == 0:
        assert o is None or len(o[0]) == 0
        assert library_load(tmp, obj_path, t, rots, s, []) is None
        assert not expect_tris
        return None
    for k in range(3):  # positions, normals, texture coordinates
        assert np.array_equal(_bits(o[k]), _bits(r[k])), ("oracle", k)
    assert np.array_equal(o[3], r[3]) and o[4] == r[4]
    lib = library_load(tmp, obj_path, t, rots, s, [m for _, m in r[4]])
    for k in range(3):
        assert np.array_equal(_bits(lib[k]), _bits(r[k])), ("library", k)
    # one library mesh per reference mesh, same triangles, bound as ResolveMaterialName binds
    mesh4, index = lib[3], lib[4]
    counts = np.bincount(r[3], minlength=len(r[4]))
    assert mesh4[:, 3].tolist() == counts.tolist()
    assert mesh4[:, 0].tolist() == [index.get(m, index["dm"]) if m else index["dm"] for _, m in r[4]]
    assert expect_tris
    return r


def _write(path, lines, eol="\n"):
    with open(path, "w", newline="") as f:
        f.write(eol.join(lines) + eol)
    return path


def test_eigen_inverse_restatement_matches_eigen():
    rng = np.random.default_rng(5)
    with np.errstate(all="ignore"):
        for it in range(400):
            m = rng.normal(0, 3, (4, 4)).astype(np.float32)
            if it % 2:
                m[3] = [0, 0, 0, 1]
            out = np.zeros(16, np.float32)
            ref().ref_inverse4(np.ascontiguousarray(m).ctypes.data, out.ctypes.data)
            assert np.array_equal(_bits(rtref_obj.inverse4(m)), _bits(out.reshape(4, 4)))


def test_object_transforms_match_eigen():
    rng = np.random.default_rng(6)
    for _ in range(200):
        t = rng.uniform(-500, 500, 3).astype(np.float32)
        s = rng.uniform(0.01, 400, 3).astype(np.float32)
        rots = [(int(rng.integers(0, 3)), float(np.float32(rng.uniform(-360, 360)))) for _ in range(rng.integers(0, 4))]
        axes = np.array([a for a, _ in rots] or [0], np.int32)
        degs = np.array([d for _, d in rots] or [0], np.float32)
        xf, inv = np.zeros(16, np.float32), np.zeros(16, np.float32)
        ref().ref_object_transform(t.ctypes.data, axes.ctypes.data, degs.ctypes.data, len(rots), s.ctypes.data,
                                   xf.ctypes.data, inv.ctypes.data)
        oxf, oinv = rtref_obj.object_transform(_obj_dict(t, rots, s))
        assert np.array_equal(_bits(oxf), _bits(xf.reshape(4, 4)))
        assert np.array_equal(_bits(oinv), _bits(inv.reshape(4, 4)))


def _mesh_file(tmp, rng, n_v=160, n_f=260, kmax=8, planar=False):
    lines = []
    if planar:  # concave planar polygons in random planes: stars and combs
        faces, vid = [], 0
        for f in range(n_f // 4):
            k = int(rng.integers(5, kmax + 1))
            ang = np.sort(rng.uniform(0, 2 * np.pi, k))
            rad = np.where(np.arange(k) % 2 == 0, 1.0, rng.uniform(0.2, 0.6))
            pts2 = np.stack([np.cos(ang) * rad, np.sin(ang) * rad], 1)
            basis = np.linalg.qr(rng.normal(size=(3, 3)))[0]
            c = rng.normal(0, 3, 3)
            for p in pts2:
                q = c + basis[:, 0] * p[0] + basis[:, 1] * p[1]
                lines.append("v %.7g %.7g %.7g" % tuple(q))
            faces.append(list(range(vid + 1, vid + k + 1)) if f % 2 else list(range(vid + k, vid, -1)))
            vid += k
        lines += ["f " + " ".join(str(i) for i in fc) for fc in faces]
        return lines
    V = rng.normal(0, 1, (n_v, 3))
    for v in V:
        lines.append("v %.6f %.6f %.6f" % tuple(v))
    for v in rng.normal(0, 1, (n_v, 3)):
        lines.append("vn %.6f %.6f %.6f" % tuple(v))
    for v in rng.uniform(0, 1, (n_v, 2)):
        lines.append("vt %.6f %.6f" % tuple(v))
    for f in range(n_f):
        k = int(rng.integers(3, kmax + 1))
        idx = rng.choice(n_v, k, replace=False) + 1
        form = f % 4
        if form == 0:
            lines.append("f " + " ".join(str(i) for i in idx))
        elif form == 1:
            lines.append("f " + " ".join(f"{i}/{i}/{i}" for i in idx))
        elif form == 2:
            lines.append("f " + " ".join(f"{i}//{i}" for i in idx))
        else:
            lines.append("f " + " ".join(f"{i}/{i}" for i in idx))
    return lines


TRANSFORMS = [((0, 0, 0), [], (1, 1, 1)), ((1.5, -2.25, 3.0), [(1, -72.9), (0, 14.0), (2, 33.3)], (1.7, 0.6, 2.3)),
              ((423.0, 0.0, 247.0), [(1, -72.9)], (166.4, 330.0, 165.4)), ((-7, 3, 1e3), [(2, 188.0)], (0.03, 5, 66))]


@pytest.mark.parametrize("ti", range(len(TRANSFORMS)))
def test_random_polygons_match_reference(tmp_path, ti):
    """Non-planar random polygons of 3-8 corners (tinyobj's ear clipping gives
    up on some, dropping their remaining corners), all corner forms."""
    rng = np.random.default_rng(10 + ti)
    p = _write(str(tmp_path / "m.obj"), _mesh_file(str(tmp_path), rng))
    r = check_all(str(tmp_path), p, *TRANSFORMS[ti])
    assert r is not None and len(r[0]) > 300


def test_concave_planar_polygons_match_reference(tmp_path):
    rng = np.random.default_rng(20)
    p = _write(str(tmp_path / "m.obj"), _mesh_file(str(tmp_path), rng, n_f=240, kmax=12, planar=True))
    r = check_all(str(tmp_path), p, *TRANSFORMS[1])
    fan = sum(len(l.split()) - 3 for l in open(p) if l.startswith("f "))
    assert r is not None and len(r[0]) > 0.9 * fan  # planar polygons clip fully


def _number_tokens(rng, n):
    toks = []
    for _ in range(n):
        kind = int(rng.integers(0, 10))
        x = rng.normal(0, 10 ** rng.uniform(-3, 4))
        if kind == 0:
            toks.append("%.20f" % x)
        elif kind == 1:
            toks.append("%.17g" % x)
        elif kind == 2:
            toks.append("%.9e" % x)
        elif kind == 3:
            toks.append(("%.6f" % x).replace("0.", ".", 1))
        elif kind == 4:
            toks.append(str(int(x * 1000)))
        elif kind == 5:
            toks.append("%.3fE%+d" % (x, int(rng.integers(-12, 12))))
        elif kind == 6:
            toks.append(rng.choice(["abc", "1e", "--1", "+", ".", "-.", "1.", "+.5e-2", "7e+", "3.5x", "2e5q", "0x10",
                                    "1e0", "-0", "00012.5000"]))
        elif kind == 7:
            toks.append("%.12f" % x)
        else:
            toks.append("%.7g" % x)
    return toks


def test_number_reader_matches_reference(tmp_path):
    """tinyobj's own digit-by-digit number reader, not a correctly rounded one:
    long fractions, exponents, leading dots and unreadable fields (which read
    as 0)."""
    rng = np.random.default_rng(30)
    toks = _number_tokens(rng, 3 * 1200)
    lines = ["v " + " ".join(toks[3 * i:3 * i + 3]) for i in range(1200)]
    lines += [f"f {i} {i + 1} {i + 2}" for i in range(1, 1199, 3)]
    p = _write(str(tmp_path / "n.obj"), lines)
    check_all(str(tmp_path), p)
    # the vectorised reader equals the one-field restatement on every token
    vec = rtref_obj.parse_reals([t.encode() for t in toks])
    one = np.array([np.float32(v if v is not None else 0.0)
                    for v in (rtref_obj.parse_double(t.encode()) for t in toks)], np.float32)
    assert np.array_equal(_bits(vec), _bits(one))


def test_shape_and_material_statements_match_reference(tmp_path):
    """o / g / usemtl / mtllib as LoadObj handles them: shapes split on o and g
    (with tinyobj's rules for empty groups), materials only from a readable MTL
    (the first readable file of an mtllib line; names with trailing blanks
    trimmed there but not in usemtl; the first definition of a name wins),
    unknown usemtl names carry no material, l and p lines, and faces lost when
    a usemtl is directly followed by o."""
    d = str(tmp_path)
    _write(os.path.join(d, "a.mtl"), ["# materials", "newmtl red  ", "Kd 1 0 0", "newmtl blue", "newmtl red",
                                      "newmtl green\t"])
    _write(os.path.join(d, "b.mtl"), ["newmtl other"])
    quad = lambda b: f"f {b} {b + 1} {b + 2} {b + 3}"
    lines = ["mtllib missing.mtl a.mtl b.mtl"]
    lines += ["v %d %d %d" % (x, y, z) for z in range(6) for x, y in ((0, 0), (1, 0), (1, 1), (0, 1))]
    lines += [quad(1), "usemtl red", quad(5), "usemtl unknown", quad(9), "usemtl blue ", quad(13),
              "o second object", "usemtl green", quad(17), "g", quad(21), "g a  b\tc", "usemtl red", quad(1),
              "l 1 2 3", "o third", "p 4", "g lone", "usemtl blue", "o fourth", quad(5), "usemtl green", quad(9),
              "usemtl red", "o fifth", "s 1", quad(13), "g", "g", quad(17), "usemtl other"]
    p = _write(os.path.join(d, "s.obj"), lines)
    r = check_all(d, p)
    # what tinyobj makes of it: "usemtl blue " (trailing blank) names no material,
    # and the faces before "o third" / "o fourth" after a usemtl change are lost
    assert [m for _, m in r[4]] == ["", "red", "green", "red", "red"]


@pytest.mark.parametrize("eol", ["\r\n", "\r"])
def test_line_endings_and_relative_indices_match_reference(tmp_path, eol):
    rng = np.random.default_rng(40)
    lines = ["o rel"]
    for k in range(40):
        lines += ["v %.6f %.6f %.6f" % tuple(rng.normal(0, 1, 3)) for _ in range(5)]
        lines += ["vt %.6f %.6f" % tuple(rng.uniform(0, 1, 2)) for _ in range(5)]
        lines += ["vn %.6f %.6f %.6f" % tuple(rng.normal(0, 1, 3)) for _ in range(5)]
        lines.append("f " + " ".join(f"{-i}/{-i}/{-i}" for i in range(5, 0, -1)))
        lines.append("f  -1  -3   -4\t")
    p = _write(str(tmp_path / "r.obj"), lines, eol)
    check_all(str(tmp_path), p, *TRANSFORMS[3])


def test_unreadable_index_makes_the_object_empty(tmp_path):
    """An index of 0 (or no number) makes tinyobj::LoadObj fail; the reference
    then skips the object, and so do the library and the oracle."""
    lines = ["v 0 0 0", "v 1 0 0", "v 0 1 0", "f 1 2 3", "f 1 0 3"]
    p = _write(str(tmp_path / "z.obj"), lines)
    check_all(str(tmp_path), p, expect_tris=False)
