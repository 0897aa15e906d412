"""OBJ import and object transforms of the oracle (TEST INFRASTRUCTURE ONLY).

A plain restatement, written independently of the library's importer, of what
the reference does to an OBJ file: OBJLoader::LoadObject
(Src/Utils/objloader.cpp:66-210) over its vendored tinyobjloader
(ThirdParty/tinyobj/tiny_obj_loader.h, LoadObj with triangulation) and the
object transform of Scene::ParseObjects (Src/Core/scene.cpp:405-448) in
Eigen 3.4.0's arithmetic.  Each function cites the code it follows.  Pinned by
tests against a build of the reference's own tinyobjloader and Eigen
(oracle/objload_ref.cpp) on generated OBJ files.

Deviations, only where the reference reads out of bounds: a corner past the
last element of the file is an error, and the ear clipping sees the vertices
read up to the face's own line.
"""
from __future__ import annotations

import ctypes
import math
import os
import re
from typing import Optional

import numpy as np

F32 = np.float32
_LIBM = ctypes.CDLL("libm.so.6")
_LIBM.cosf.argtypes = [ctypes.c_float]
_LIBM.cosf.restype = ctypes.c_float
_LIBM.sinf.argtypes = [ctypes.c_float]
_LIBM.sinf.restype = ctypes.c_float


# ------------------------------------------------------------------ transforms
def deg2rad(a) -> np.float32:
    """AngleToRadian (Src/Utils/mathutil.cu:4-7): angle / 180 in float, times the
    double PI = 3.1415926 (Src/Core/global.h:7), rounded to float."""
    return F32(float(F32(a) / F32(180.0)) * 3.1415926)


def rotation(axis: int, deg) -> np.ndarray:
    """RotateX / RotateY / RotateZ (mathutil.cu:107-135); cos and sin of a float
    are the C library's float functions."""
    rad = deg2rad(deg)
    c, s = F32(_LIBM.cosf(float(rad))), F32(_LIBM.sinf(float(rad)))
    R = np.eye(4, dtype=np.float32)
    if axis == 0:
        R[1, 1], R[1, 2], R[2, 1], R[2, 2] = c, -s, s, c
    elif axis == 1:
        R[0, 0], R[0, 2], R[2, 0], R[2, 2] = c, s, -s, c
    else:
        R[0, 0], R[0, 1], R[1, 0], R[1, 1] = c, -s, s, c
    return R


def mat4_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Eigen's 4x4 float product: entry (i, j) = ((a_i0 b_0j + a_i1 b_1j) + a_i2 b_2j) + a_i3 b_3j."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            r[i, j] = ((a[i, 0] * b[0, j] + a[i, 1] * b[1, j]) + a[i, 2] * b[2, j]) + a[i, 3] * b[3, j]
    return r


def inverse4(m: np.ndarray) -> np.ndarray:
    """Matrix4f::inverse as Eigen 3.4.0 computes it with SSE
    (Eigen/src/LU/arch/InverseSize4.h): the adjugate over the determinant, both
    formed from the four 2x2 blocks, every product and sum in its order."""
    g = lambda r, c: F32(m[r, c])
    # blocks read column by column: A = rows 0-1 x cols 0-1, B = rows 2-3 x cols 0-1,
    # C = rows 0-1 x cols 2-3, D = rows 2-3 x cols 2-3
    a = [g(0, 0), g(1, 0), g(0, 1), g(1, 1)]
    b = [g(2, 0), g(3, 0), g(2, 1), g(3, 1)]
    c = [g(0, 2), g(1, 2), g(0, 3), g(1, 3)]
    d = [g(2, 2), g(3, 2), g(2, 3), g(3, 3)]
    ab = [a[3] * b[0] - a[1] * b[2], a[3] * b[1] - a[1] * b[3], a[0] * b[2] - a[2] * b[0], a[0] * b[3] - a[2] * b[1]]
    dc = [d[3] * c[0] - d[1] * c[2], d[3] * c[1] - d[1] * c[3], d[0] * c[2] - d[2] * c[0], d[0] * c[3] - d[2] * c[1]]
    da, db = a[3] * a[0] - a[1] * a[2], b[3] * b[0] - b[1] * b[2]
    dcc, dd = c[3] * c[0] - c[1] * c[2], d[3] * d[0] - d[1] * d[2]
    tr = (dc[0] * ab[0] + dc[1] * ab[2]) + (dc[2] * ab[1] + dc[3] * ab[3])
    det = (da * dd + db * dcc) - tr
    if det == 0 or not np.isfinite(det):
        raise ValueError("object transform is singular")
    rd = F32(1.0) / det
    sg = [rd, -rd, -rd, rd]
    ia = [a[0] * dd - (b[0] * dc[0] + b[1] * dc[2]), a[1] * dd - (b[0] * dc[1] + b[1] * dc[3]),
          a[2] * dd - (b[2] * dc[0] + b[3] * dc[2]), a[3] * dd - (b[2] * dc[1] + b[3] * dc[3])]
    idd = [d[0] * da - (c[0] * ab[0] + c[1] * ab[2]), d[1] * da - (c[0] * ab[1] + c[1] * ab[3]),
           d[2] * da - (c[2] * ab[0] + c[3] * ab[2]), d[3] * da - (c[2] * ab[1] + c[3] * ab[3])]
    ib = [c[0] * db - (d[0] * ab[3] - d[1] * ab[2]), c[1] * db - (d[1] * ab[0] - d[0] * ab[1]),
          c[2] * db - (d[2] * ab[3] - d[3] * ab[2]), c[3] * db - (d[3] * ab[0] - d[2] * ab[1])]
    ic = [b[0] * dcc - (a[0] * dc[3] - a[1] * dc[2]), b[1] * dcc - (a[1] * dc[0] - a[0] * dc[1]),
          b[2] * dcc - (a[2] * dc[3] - a[3] * dc[2]), b[3] * dcc - (a[3] * dc[0] - a[2] * dc[1])]
    ia, ib, ic, idd = ([x[k] * sg[k] for k in range(4)] for x in (ia, ib, ic, idd))
    r = np.zeros((4, 4), np.float32)
    r[:, 0] = [ia[3], ia[1], ib[3], ib[1]]
    r[:, 1] = [ia[2], ia[0], ib[2], ib[0]]
    r[:, 2] = [ic[3], ic[1], idd[3], idd[1]]
    r[:, 3] = [ic[2], ic[0], idd[2], idd[0]]
    return r


def object_transform(o: dict):
    """transform = T * R * S and its inverse (Scene::ParseObjects, scene.cpp:405-448)."""
    tr, rot, scl = (np.eye(4, dtype=np.float32) for _ in range(3))
    if "translate" in o:
        tr[0:3, 3] = [F32(x) for x in o["translate"]]
    for r in o.get("rotate", []):
        for ax, key in enumerate(("rotateX", "rotateY", "rotateZ")):
            if key in r:
                rot = mat4_mul(rotation(ax, F32(r[key])), rot)
                break
    if "scale" in o:
        scl[0, 0], scl[1, 1], scl[2, 2] = (F32(x) for x in o["scale"])
    xf = mat4_mul(mat4_mul(tr, rot), scl)
    return xf, inverse4(xf)


def xform_points(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """transform * (p, 1): Eigen sums a column-major product column by column."""
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = ((m[r, 0] * p[:, 0] + m[r, 1] * p[:, 1]) + m[r, 2] * p[:, 2]) + m[r, 3] * F32(1)
    return out


def _dot3(a, b):
    """Eigen's 3-term reduction a0 + (a1 + a2)."""
    return a[..., 0] * b[..., 0] + (a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2])


def normalize(v: np.ndarray) -> np.ndarray:
    """normalized(): v / sqrt(squaredNorm), unchanged when the norm is 0."""
    with np.errstate(over="ignore"):
        z = _dot3(v, v)
    s = np.sqrt(z)
    out = v.copy()
    m = z > 0
    out[m] = v[m] / s[m][:, None]
    return out


def xform_normals(inv: np.ndarray, n: np.ndarray) -> np.ndarray:
    """(transformInv.transpose() * (n, 1)).normalized(): a row of the transposed
    (row-major) matrix times the vector is one 4-wide product summed pairwise,
    (x0 + x2) + (x1 + x3)."""
    out = np.empty_like(n)
    for r in range(3):
        out[:, r] = (inv[0, r] * n[:, 0] + inv[2, r] * n[:, 2]) + (inv[1, r] * n[:, 1] + inv[3, r] * F32(1))
    return normalize(out)


def cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


# ------------------------------------------------------------------ numbers
_FRAC = [1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001]


def parse_double(tok: bytes) -> Optional[float]:
    """tinyobj's tryParseDouble (tiny_obj_loader.h:804-930) on one field; None
    where it fails."""
    n, i = len(tok), 0
    if n == 0:
        return None
    mant, expo, sign, lead_dot = 0.0, 0, 1, False
    isd = lambda k: k < n and 48 <= tok[k] <= 57
    if tok[0] in b"+-":
        sign = -1 if tok[0] == 45 else 1
        i = 1
        lead_dot = i < n and tok[i] == 46
    elif isd(0):
        pass
    elif tok[0] == 46:
        lead_dot = True
    else:
        return None
    if not lead_dot:
        read = 0
        while isd(i):
            mant = mant * 10 + (tok[i] - 48)
            i += 1
            read += 1
        if read == 0:
            return None
    if i >= n:
        return sign * mant
    if tok[i] == 46:
        i += 1
        read = 1
        while isd(i):
            mant += (tok[i] - 48) * (_FRAC[read] if read < 8 else math.pow(10.0, -read))
            read += 1
            i += 1
    elif tok[i] not in b"eE":
        return sign * mant
    if i >= n:
        return sign * mant
    if tok[i] in b"eE":
        i += 1
        esign = 1
        if i < n and tok[i] in b"+-":
            esign = -1 if tok[i] == 45 else 1
            i += 1
        elif not isd(i):
            return None
        read = 0
        while isd(i):
            expo = expo * 10 + (tok[i] - 48)
            i += 1
            read += 1
        expo *= esign
        if read == 0:
            return None
    return sign * (math.ldexp(mant * math.pow(5.0, expo), expo) if expo else mant)


def _scales(n: int) -> np.ndarray:
    """The factor tryParseDouble gives fraction digit k: the table for k < 8, pow beyond."""
    return np.array([_FRAC[k] if k < 8 else math.pow(10.0, -k) for k in range(n)], np.float64)


def parse_reals(tokens: list, default: float = 0.0) -> np.ndarray:
    """parseReal over many fields, as float32, in batches of 2^20 fields."""
    step = 1 << 20
    parts = [_parse_reals(tokens[i:i + step], default) for i in range(0, len(tokens), step)]
    return np.concatenate(parts) if parts else np.zeros(0, np.float32)


def _parse_reals(tokens: list, default: float = 0.0) -> np.ndarray:
    """parseReal over many fields at once, as float32.  Fields of the plain forms
    [sign] digits [. digits] [e [sign] digits] and [sign] . digits [e ...] (at
    most 32 characters) are assembled column by column with numpy in
    tryParseDouble's order (integer digits m = m * 10 + d, fraction digit k
    m = m + d * scale_k, then the exponent); any other field goes through
    parse_double one at a time."""
    n = len(tokens)
    out = np.full(n, default, np.float64)
    if n == 0:
        return out.astype(np.float32)
    W = max(1, min(32, max(len(t) for t in tokens)))
    arr = np.array(tokens, dtype=f"S{W + 1}")
    A = arr.view(np.uint8).reshape(n, W + 1)[:, :W]
    length = np.char.str_len(arr)
    col = np.arange(W)[None, :]
    inside = col < length[:, None]
    dig = (A >= 48) & (A <= 57) & inside
    dot = (A == 46) & inside
    ee = ((A == 101) | (A == 69)) & inside
    sg = ((A == 43) | (A == 45)) & inside
    n_dot, n_e = dot.sum(1), ee.sum(1)
    start = sg[:, 0].astype(np.int64)
    epos = np.where(n_e > 0, ee.argmax(1), length)
    dpos = np.where(n_dot > 0, dot.argmax(1), -1)
    esign = (n_e > 0) & (np.take_along_axis(sg, np.minimum(epos + 1, W - 1)[:, None], 1)[:, 0]) & (epos + 1 < length)
    # what may stand in each column of a plain field
    mant_end = np.where(dpos >= 0, dpos, epos)
    is_int = dig & (col >= start[:, None]) & (col < mant_end[:, None])
    is_frac = dig & (col > dpos[:, None]) & (col < epos[:, None]) & (dpos[:, None] >= 0)
    exp_from = epos + 1 + esign
    is_exp = dig & (col >= exp_from[:, None]) & (n_e > 0)[:, None]
    ok = (n_dot <= 1) & (n_e <= 1) & (np.array([len(t) for t in tokens]) <= W) & (length > 0)
    ok &= (sg.sum(1) == start + esign)  # signs only in front and after the 'e'
    ok &= (is_int | is_frac | is_exp | dot | ee | sg).sum(1) == length
    ok &= (dpos < 0) | (dpos < epos)
    n_int, n_exp = is_int.sum(1), is_exp.sum(1)
    ok &= (n_int > 0) | (dpos == start)  # digits before the '.', or a leading '.'
    ok &= (n_e == 0) | (n_exp > 0)
    ok &= (n_e == 0) | (n_exp == length - exp_from)
    d = A.astype(np.float64) - 48
    mant = np.zeros(n, np.float64)
    for j in range(W):
        if is_int[:, j].any():
            mant = np.where(is_int[:, j], mant * 10 + d[:, j], mant)
    sc = _scales(W + 1)
    k = col - dpos[:, None]
    for j in range(W):
        if is_frac[:, j].any():
            mant = np.where(is_frac[:, j], mant + d[:, j] * sc[np.clip(k[:, j], 0, W)], mant)
    expo = np.zeros(n, np.int64)
    for j in range(W):
        if is_exp[:, j].any():
            expo = np.where(is_exp[:, j], expo * 10 + (A[:, j].astype(np.int64) - 48), expo)
    eneg = esign & (np.take_along_axis(A, np.minimum(epos + 1, W - 1)[:, None], 1)[:, 0] == 45)
    expo = np.where(eneg, -expo, expo)
    val = mant.copy()
    for e in np.unique(expo[ok & (expo != 0)]):
        sel = ok & (expo == e)
        val[sel] = np.ldexp(mant[sel] * math.pow(5.0, int(e)), int(e))
    neg = sg[:, 0] & (A[:, 0] == 45)
    val = np.where(neg, -1 * val, val)
    out[ok] = val[ok]
    for i in np.flatnonzero(~ok):
        v = parse_double(tokens[i])
        out[i] = default if v is None else v
    with np.errstate(over="ignore"):
        return out.astype(np.float32)


# ------------------------------------------------------------------ triangulation
def _pnpoly3(vx, vy, tx, ty) -> bool:
    """pnpoly (tiny_obj_loader.h:1316-1327), float32."""
    c = False
    j = 2
    for i in range(3):
        if ((vy[i] > ty) != (vy[j] > ty)) and (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]):
            c = not c
        j = i
    return c


def ear_clip(vi: list, V: np.ndarray, nv: int) -> list:
    """tinyobj's triangulation of one polygon (exportGroupsToShape,
    tiny_obj_loader.h:1358-1560) in float32.  vi: the corners' vertex indices;
    V: vertex coordinates (float32, 3 per vertex); nv: the vertices it may read.
    Returns triangles as corner positions."""
    npl = len(vi)
    ok = lambda v: 0 <= v < nv
    ax0, ax1 = 1, 2
    eps = F32(np.finfo(np.float32).eps)
    for k in range(npl):
        a, b, c = vi[k % npl], vi[(k + 1) % npl], vi[(k + 2) % npl]
        if not (ok(a) and ok(b) and ok(c)):
            continue
        p0, p1, p2 = V[a], V[b], V[c]
        e0 = p1 - p0
        e1 = p2 - p1
        cx = abs(e0[1] * e1[2] - e0[2] * e1[1])
        cy = abs(e0[2] * e1[0] - e0[0] * e1[2])
        cz = abs(e0[0] * e1[1] - e0[1] * e1[0])
        if cx > eps or cy > eps or cz > eps:
            if not (cx > cy and cx > cz):
                ax0 = 0
                if cz > cx and cz > cy:
                    ax1 = 1
            break
    area = F32(0)
    for k in range(npl):
        a, b = vi[k % npl], vi[(k + 1) % npl]
        if not (ok(a) and ok(b)):
            continue
        area += (V[a][ax0] * V[b][ax1] - V[a][ax1] * V[b][ax0]) * F32(0.5)
    rem = list(range(npl))
    guess, iters, prev = 0, npl, npl
    tris = []
    while len(rem) > 3 and iters > 0:
        n = len(rem)
        if guess >= n:
            guess -= n
        if prev != n:
            prev, iters = n, n
        else:
            iters -= 1
        ind = [rem[(guess + k) % n] for k in range(3)]
        vx = [V[vi[i]][ax0] if ok(vi[i]) else F32(0) for i in ind]
        vy = [V[vi[i]][ax1] if ok(vi[i]) else F32(0) for i in ind]
        cr = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1])
        if cr * area < 0:
            guess += 1
            continue
        overlap = False
        for other in range(3, n):
            v = vi[rem[(guess + other) % n]]
            if ok(v) and _pnpoly3(vx, vy, V[v][ax0], V[v][ax1]):
                overlap = True
                break
        if overlap:
            guess += 1
            continue
        tris.append(tuple(ind))
        del rem[(guess + 1) % n]
    if len(rem) == 3:
        tris.append(tuple(rem))
    return tris


# ------------------------------------------------------------------ files
def _lines(data: bytes) -> list:
    """safeGetline: lines end at \\n, \\r or \\r\\n."""
    return data.splitlines()


def _mtl_names(path: str) -> Optional[list]:
    """'newmtl' names of an MTL file in order (LoadMtl, tiny_obj_loader.h:1638-2012),
    the last material flushed even without a name; None when it cannot be opened."""
    try:
        with open(path, "rb") as f:
            data = f.read()
    except OSError:
        return None
    names, cur = [], b""
    for line in _lines(data):
        line = line.rstrip(b" \t")
        t = line.lstrip(b" \t")
        if not t or t[:1] == b"#":
            continue
        if t[:6] == b"newmtl" and t[6:7] in (b" ", b"\t"):
            if cur:
                names.append(cur)
            cur = t[7:]
        # other statements carry no name
    names.append(cur)
    return [n.decode("latin-1") for n in names]


def _atoi(t: bytes, i: int) -> int:
    n = len(t)
    while i < n and t[i] in b" \t\n\v\f\r":
        i += 1
    neg = False
    if i < n and t[i] in b"+-":
        neg = t[i] == 45
        i += 1
    v = 0
    while i < n and 48 <= t[i] <= 57:
        v = v * 10 + (t[i] - 48)
        i += 1
    return -v if neg else v


def _skip_to(t: bytes, i: int, stops: bytes) -> int:
    n = len(t)
    while i < n and t[i] not in stops:
        i += 1
    return i


_FIELDS = re.compile(rb"[^ \t\r]+")
def _fix(k: int, n: int) -> int:
    """fixIndex (tiny_obj_loader.h:738-759)."""
    if k == 0:
        raise _Unreadable()
    return k - 1 if k > 0 else n + k


def _face_slow(t: bytes, nv: int, nvt: int, nvn: int):
    """parseTriple over a face line (tiny_obj_loader.h:1070-1120, 2270-2300), character by character."""
    i = 2
    while i < len(t) and t[i] in (32, 9):
        i += 1
    vs, ts, ns = [], [], []
    while i < len(t) and t[i] not in (13, 10):
        v = _fix(_atoi(t, i), nv)
        i = _skip_to(t, i, b"/ \t\r")
        vt = vn = -1
        if i < len(t) and t[i] == 47:
            i += 1
            if i < len(t) and t[i] == 47:
                i += 1
                vn = _fix(_atoi(t, i), nvn)
                i = _skip_to(t, i, b"/ \t\r")
            else:
                vt = _fix(_atoi(t, i), nvt)
                i = _skip_to(t, i, b"/ \t\r")
                if i < len(t) and t[i] == 47:
                    i += 1
                    vn = _fix(_atoi(t, i), nvn)
                    i = _skip_to(t, i, b"/ \t\r")
        vs.append(v)
        ts.append(vt)
        ns.append(vn)
        while i < len(t) and t[i] in (32, 9, 13):
            i += 1
    return vs, ts, ns


def _faces_bulk(lines: list, counts: np.ndarray):
    """Corners of many face lines at once.  lines: the text after "f "; counts:
    (n, 3) element counts (v, vt, vn) at each line.  Returns per-face corner
    counts and, per corner, resolved (v, vt, vn) indices (-1 absent), with the
    lines that are not all plain v, v/vt, v//vn or v/vt/vn numbers read by
    parseTriple character by character.  Raises _Unreadable where LoadObj fails."""
    n = len(lines)
    tok = [ln.split(b" ") if b"\t" not in ln and b"\x0b" not in ln and b"\x0c" not in ln else None for ln in lines]
    slow = np.array([t is None for t in tok], bool)
    tok = [[c for c in t if c] if t is not None else [] for t in tok]
    ncor = np.fromiter((len(t) for t in tok), np.int64, n)
    flat = [c for t in tok for c in t]
    first = np.concatenate([[0], np.cumsum(ncor)[:-1]]) if n else np.zeros(0, np.int64)
    cv = np.full(len(flat), -1, np.int64)
    ct = np.full(len(flat), -1, np.int64)
    cn = np.full(len(flat), -1, np.int64)
    if flat:
        arr = np.array(flat, dtype="S48")
        nslash = np.char.count(arr, b"/")
        dbl = np.char.find(arr, b"//") >= 0
        p0 = np.char.partition(arr, b"/")
        v_s, rest = p0[:, 0], p0[:, 2]
        p1 = np.char.partition(rest, b"/")
        a_s, b_s = p1[:, 0], p1[:, 2]

        def plain(x):  # [+-]digits
            body = np.char.lstrip(x, b"+-")
            return (np.char.str_len(body) > 0) & np.char.isdigit(body) & (np.char.str_len(x) - np.char.str_len(body) <= 1)

        ok = plain(v_s) & (nslash <= 2)
        ok &= (nslash == 0) | ((nslash == 1) & plain(a_s))
        ok2 = (nslash == 2) & plain(v_s) & ((dbl & (a_s == b"") & plain(b_s)) | (~dbl & plain(a_s) & plain(b_s)))
        ok = (ok & (nslash <= 1)) | ok2
        line_of = np.repeat(np.arange(n), ncor)
        bad_line = np.zeros(n, bool)
        np.logical_or.at(bad_line, line_of[~ok], True)
        slow |= bad_line
        good = ok & ~slow[line_of]

        def ints(x, sel):
            out = np.zeros(len(x), np.int64)
            out[sel] = x[sel].astype(np.int64)
            return out

        rv = ints(v_s, good)
        rt_ = ints(a_s, good & (nslash >= 1) & (a_s != b""))
        rn = ints(np.where(nslash == 2, b_s, b""), good & (nslash == 2))
        has_t = good & (nslash >= 1) & (a_s != b"")
        has_n = good & (nslash == 2)
        if ((rv == 0) & good).any() or ((rt_ == 0) & has_t).any() or ((rn == 0) & has_n).any():
            raise _Unreadable()
        cnt = counts[line_of]
        fix = lambda r, c: np.where(r > 0, r - 1, c + r)
        cv[good] = fix(rv, cnt[:, 0])[good]
        ct[has_t] = fix(rt_, cnt[:, 1])[has_t]
        cn[has_n] = fix(rn, cnt[:, 2])[has_n]
    if slow.any():  # rebuild with the slow lines read one at a time
        parts_n, parts_v, parts_t, parts_n2 = [], [], [], []
        for i in range(n):
            if slow[i]:
                vs, ts, ns = _face_slow(b"f " + lines[i], *counts[i])
            else:
                sl = slice(first[i], first[i] + ncor[i])
                vs, ts, ns = cv[sl], ct[sl], cn[sl]
            parts_n.append(len(vs))
            parts_v.append(np.asarray(vs, np.int64))
            parts_t.append(np.asarray(ts, np.int64))
            parts_n2.append(np.asarray(ns, np.int64))
        ncor = np.array(parts_n, np.int64)
        cv, ct, cn = (np.concatenate(p) if p else np.zeros(0, np.int64) for p in (parts_v, parts_t, parts_n2))
    return ncor, cv, ct, cn


def load_obj(path: str, xf: np.ndarray, xf_inv: np.ndarray):
    """OBJLoader::LoadObject (objloader.cpp:66-210) over tinyobj::LoadObj: a list
    of meshes {name, material, pos (n,9), nrm (n,9), uv (n,6)}; [] when LoadObj fails."""
    with open(path, "rb") as f:
        data = f.read()
    vtok, vntok, vttok = [], [], []
    face_lines, face_counts = [], []
    stream = []    # file order: ("f", first face, end face) runs and (kind, text) statements
    fields = _FIELDS.findall
    run_start = None
    for raw in _lines(data):
        t = raw.lstrip(b" \t")
        if not t or t[0] == 35:  # '#'
            continue
        c0, c1 = t[0], t[1:2]
        if c0 == 102 and c1 in (b" ", b"\t"):  # f
            face_lines.append(t[2:])
            face_counts.append((len(vtok) // 3, len(vttok) // 2, len(vntok) // 3))
            if run_start is None:
                run_start = len(face_lines) - 1
            continue
        if c0 == 118 and c1 in (b" ", b"\t"):  # v
            f = fields(t, 2)[:3]
            vtok.extend(f + [b""] * (3 - len(f)))
            continue
        if c0 == 118 and t[1:2] == b"n" and t[2:3] in (b" ", b"\t"):
            f = fields(t, 3)[:3]
            vntok.extend(f + [b""] * (3 - len(f)))
            continue
        if c0 == 118 and t[1:2] == b"t" and t[2:3] in (b" ", b"\t"):
            f = fields(t, 3)[:2]
            vttok.extend(f + [b""] * (2 - len(f)))
            continue
        ev = None
        if c0 in (108, 112) and c1 in (b" ", b"\t"):  # l, p
            ev = ("prim", None)
        elif c0 == 103 and c1 in (b" ", b"\t"):  # g
            ev = ("g", b" ".join(fields(t)[1:]).decode("latin-1"))
        elif c0 == 111 and c1 in (b" ", b"\t"):  # o
            ev = ("o", t[2:].decode("latin-1"))
        elif t[:6] == b"usemtl" and t[6:7] in (b" ", b"\t"):
            ev = ("usemtl", t[7:].decode("latin-1"))
        elif t[:6] == b"mtllib" and t[6:7] in (b" ", b"\t"):
            ev = ("mtllib", t[7:].decode("latin-1"))
        if ev is not None:
            if run_start is not None:
                stream.append(("f", run_start, len(face_lines)))
                run_start = None
            stream.append(ev)
    if run_start is not None:
        stream.append(("f", run_start, len(face_lines)))
    try:
        ncor, cv, ct, cn = _faces_bulk(face_lines, np.array(face_counts, np.int64).reshape(-1, 3))
    except _Unreadable:
        return []
    V = parse_reals(vtok).reshape(-1, 3)
    VN = parse_reals(vntok).reshape(-1, 3)
    VT = parse_reals(vttok).reshape(-1, 2)
    nfaces = len(face_lines)
    first = np.concatenate([[0], np.cumsum(ncor)[:-1]]).astype(np.int64) if nfaces else np.zeros(0, np.int64)
    # triangles per face: a triangle passes through, polygons are ear-clipped
    poly_tris = {}
    ntri = np.where(ncor == 3, 1, 0)
    for fi in np.flatnonzero(ncor > 3):
        tr = ear_clip(cv[first[fi]:first[fi] + ncor[fi]].tolist(), V, int(face_counts[fi][0]))
        poly_tris[int(fi)] = tr
        ntri[fi] = len(tr)
    cum_tri = np.concatenate([[0], np.cumsum(ntri)])
    # LoadObj's statements in order (tiny_obj_loader.h:2306-2420, 2575-2585); faces move in runs
    shapes, shape = [], {"name": "", "runs": [], "tris": False}
    name, material, group, prims = "", -1, [], False
    mat_id, mat_names = {}, []
    base = os.path.dirname(path)
    base = base + "/" if base else ""

    def export():
        if not group and not prims:
            return False
        shape["name"] = name
        for a, b in group:
            shape["runs"].append((a, b, material))
            shape["tris"] |= bool(cum_tri[b] > cum_tri[a])
        return True

    for ev in stream:
        kind = ev[0]
        if kind == "f":
            group.append((ev[1], ev[2]))
        elif kind == "prim":
            prims = True
        elif kind == "usemtl":
            mid = mat_id.get(ev[1], -1)
            if mid != material:
                export()
                group = []
                material = mid
        elif kind == "mtllib":
            for fn in ev[1].split(" "):
                names = _mtl_names(base + fn)
                if names is None:
                    continue
                for nm in names:
                    mat_id.setdefault(nm, len(mat_names))
                    mat_names.append(nm)
                break
        elif kind == "g":
            export()
            if shape["tris"]:
                shapes.append(shape)
            shape = {"name": "", "runs": [], "tris": False}
            group, prims, name = [], False, ev[1]
        elif kind == "o":
            if export():
                shapes.append(shape)
            shape = {"name": "", "runs": [], "tris": False}
            group, prims, name = [], False, ev[1]
    if export() or shape["tris"]:
        shapes.append(shape)
    # meshes per (shape, material) in order of first appearance (objloader.cpp:146-202)
    stem = os.path.basename(path)
    if "." in stem and stem.rfind(".") > 0:
        stem = stem[:stem.rfind(".")]
    meshes, order = {}, []
    for si, sh in enumerate(shapes):
        for a, b, mid in sh["runs"]:
            if cum_tri[b] == cum_tri[a]:
                continue
            key = (si, mid)
            if key not in meshes:
                mname = mat_names[mid] if 0 <= mid < len(mat_names) else ""
                sname = sh["name"] if sh["name"] else f"{stem}_shape_{si}"
                meshes[key] = {"name": sname + ("::" + mname if mname else ""), "material": mname, "c": []}
                order.append(key)
            fr = np.arange(a, b)
            tri_faces = fr[ncor[fr] == 3]
            if len(tri_faces) == b - a:  # all triangles: corners 0, 1, 2 of each face
                idx = first[fr][:, None] + np.arange(3)[None, :]
            else:
                rows = []
                for f in fr:
                    if ncor[f] == 3:
                        rows.append(first[f] + np.arange(3))
                    elif ncor[f] > 3:
                        rows.extend(first[f] + np.array(t) for t in poly_tris[int(f)])
                idx = np.array(rows, np.int64).reshape(-1, 3)
            meshes[key]["c"].append(idx)
    out = []
    for key in order:
        m = meshes[key]
        idx = np.concatenate(m.pop("c"))
        vi, ti, ni = cv[idx], ct[idx], cn[idx]
        for arr, a_ in ((vi, V), (ti, VT), (ni, VN)):
            if (arr >= len(a_)).any():
                raise ValueError(f"OBJ face references an element past the end of {path}")
        n = idx.shape[0]
        p = np.zeros((n, 3, 3), np.float32)
        has_v = vi >= 0
        p[has_v] = xform_points(xf, V[vi[has_v]])
        nrm = np.zeros((n, 3, 3), np.float32)
        has_n = ni >= 0
        if has_n.any():
            nrm[has_n] = xform_normals(xf_inv, VN[ni[has_n]])
        uv = np.zeros((n, 3, 2), np.float32)
        has_t = ti >= 0
        uv[has_t] = VT[ti[has_t]]
        face_n = ~has_n.all(axis=1)
        if face_n.any():
            fn = normalize(cross(p[face_n, 1] - p[face_n, 0], p[face_n, 2] - p[face_n, 0]))
            nrm[face_n] = fn[:, None, :]
        m["pos"] = p.reshape(n, 9)
        m["nrm"] = nrm.reshape(n, 9)
        m["uv"] = uv.reshape(n, 6)
        out.append(m)
    return out


class _Unreadable(Exception):
    """A face index tinyobj cannot read (0 or no number): LoadObj fails."""
