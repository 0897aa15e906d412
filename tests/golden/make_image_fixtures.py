"""Write the image-decoding fixtures under tests/golden/images/.

Inputs: small PNG, JPEG and PNM files covering the formats' variants.
 - PNG: every colour type and bit depth, palettes with partial alpha, tRNS keys,
   Adam7 interlacing, all five row filters, split image data.
 - JPEG: Pillow-written baseline, 12-/16-bit-table, progressive, restart-marker,
   grey, CMYK and RGB files.  Plus this script's own minimal baseline encoder,
   for sampling factors Pillow does not write (1x2, 4x1, 3x1, 2x4, 4x4).
 - PNM: P5 and P6 at 8 and 16 bits, with comments.
Expected outputs: the texels the reference's loader returns for each file.
The reference's vendored stb_image 2.30 is built from the reference tree
(`make -C oracle ref`) and called exactly as Texture::Load calls it
(Src/Core/texture.cu:207-214, four requested channels).

Run where the reference tree is present:
    python tests/golden/make_image_fixtures.py
It rewrites images/*.{png,jpg,pgm,ppm} and images/expected.npz (one uint8
array of shape (h, w, 4) per file name).
"""
from __future__ import annotations

import io
import os
import struct
import subprocess
import sys
import tempfile
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "images")
STB = os.path.join(ROOT, "oracle", "_ref", "stb_decode")


# ------------------------------------------------------------------ PNG writer
def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_rows(rows: list, bpp: int, rng) -> bytes:
    out = bytearray()
    prev = bytes(len(rows[0])) if rows else b""
    for r in rows:
        ft = int(rng.integers(0, 5))
        f = bytearray(len(r))
        for i in range(len(r)):
            a = r[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            f[i] = (r[i] - pred) & 255
        out.append(ft)
        out += f
        prev = r
    return bytes(out)


def _pack_row(samples: np.ndarray, depth: int) -> bytes:
    s = samples.astype(np.uint32).ravel()
    if depth == 16:
        return s.astype(">u2").tobytes()
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    per = 8 // depth
    out = bytearray((len(s) + per - 1) // per)
    for i, v in enumerate(s):
        out[i // per] |= int(v) << (8 - depth - (i % per) * depth)
    return bytes(out)


def write_png(path, samples, color, depth, rng, interlace=False, palette=None, trns=None, split=1):
    """samples: (h, w, channels) integers at the given depth."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    if interlace:
        passes = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
    else:
        passes = [(0, 0, 1, 1)]
    raw = bytearray()
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        raw += _filter_rows([_pack_row(sub[y], depth) for y in range(sub.shape[0])], bpp, rng)
    z = zlib.compress(bytes(raw), 9)
    parts = [z[i * len(z) // split:(i + 1) * len(z) // split] for i in range(split)]
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, int(interlace)))
    data += _chunk(b"tEXt", b"Comment\x00fixture")
    if palette is not None:
        data += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    if trns is not None:
        data += _chunk(b"tRNS", trns)
    for p in parts:
        data += _chunk(b"IDAT", p)
    data += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def png_cases(rng):
    cases = []

    def img(h, w, ch, depth):
        return rng.integers(0, 1 << depth, (h, w, ch))

    for depth in (1, 2, 4, 8, 16):
        cases.append((f"png_grey{depth}.png", dict(samples=img(13, 19, 1, depth), color=0, depth=depth)))
    g = img(11, 17, 1, 4)
    g[2:5, 3:9] = 5
    cases.append(("png_grey4_trns.png", dict(samples=g, color=0, depth=4, trns=struct.pack(">H", 5))))
    g = img(9, 10, 1, 16)
    g[::2, ::3] = 40000
    cases.append(("png_grey16_trns.png", dict(samples=g, color=0, depth=16, trns=struct.pack(">H", 40000))))
    for depth in (8, 16):
        cases.append((f"png_rgb{depth}.png", dict(samples=img(15, 21, 3, depth), color=2, depth=depth)))
        cases.append((f"png_greyalpha{depth}.png", dict(samples=img(7, 23, 2, depth), color=4, depth=depth)))
        cases.append((f"png_rgba{depth}.png", dict(samples=img(12, 9, 4, depth), color=6, depth=depth)))
    c = img(10, 14, 3, 8)
    c[1:4, 2:7] = (10, 200, 30)
    cases.append(("png_rgb8_trns.png", dict(samples=c, color=2, depth=8, trns=struct.pack(">HHH", 10, 200, 30))))
    for depth in (1, 2, 4, 8):
        n = min(1 << depth, 200)
        pal = rng.integers(0, 256, (n, 3))
        trns = bytes(rng.integers(0, 256, max(1, n // 2)).astype(np.uint8)) if depth >= 2 else None
        cases.append((f"png_palette{depth}.png",
                      dict(samples=rng.integers(0, n, (16, 11, 1)), color=3, depth=depth, palette=pal, trns=trns)))
    for name, color, ch, depth, h, w in (("png_adam7_rgba8.png", 6, 4, 8, 29, 31), ("png_adam7_grey2.png", 0, 1, 2, 9, 13),
                                         ("png_adam7_rgb16.png", 2, 3, 16, 5, 3), ("png_adam7_tiny.png", 2, 3, 8, 1, 1),
                                         ("png_adam7_palette4.png", 3, 1, 4, 17, 6)):
        kw = dict(samples=img(h, w, ch, depth) if color != 3 else rng.integers(0, 16, (h, w, 1)), color=color,
                  depth=depth, interlace=True)
        if color == 3:
            kw["palette"] = rng.integers(0, 256, (16, 3))
        cases.append((name, kw))
    cases.append(("png_split_idat.png", dict(samples=img(40, 33, 3, 8), color=2, depth=8, split=5)))
    cases.append(("png_1x1.png", dict(samples=img(1, 1, 4, 8), color=6, depth=8)))
    return cases


# ------------------------------------------------------------- JPEG encoders
def _smooth(h, w, ch, rng):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.empty((h, w, ch))
    for k in range(ch):
        fx, fy, ph = rng.uniform(0.05, 0.6), rng.uniform(0.05, 0.6), rng.uniform(0, 6)
        out[..., k] = 128 + 90 * np.sin(fx * x + ph) * np.cos(fy * y) + rng.normal(0, 18, (h, w))
    return np.clip(out, 0, 255).astype(np.uint8)


def pillow_jpeg_cases(rng):
    from PIL import Image

    cases = []

    def save(name, arr, mode=None, **kw):
        im = Image.fromarray(arr) if mode is None else Image.fromarray(arr).convert(mode)
        b = io.BytesIO()
        im.save(b, "JPEG", **kw)
        cases.append((name, b.getvalue()))

    save("jpg_444_q95.jpg", _smooth(23, 37, 3, rng), quality=95, subsampling=0)
    save("jpg_422_q75.jpg", _smooth(29, 41, 3, rng), quality=75, subsampling=1)
    save("jpg_420_q50.jpg", _smooth(35, 27, 3, rng), quality=50, subsampling=2)
    save("jpg_420_q10_large_coefs.jpg", _smooth(16, 16, 3, rng), quality=10, subsampling=2)
    save("jpg_420_1x1.jpg", _smooth(1, 1, 3, rng), quality=80, subsampling=2)
    save("jpg_422_width1.jpg", _smooth(9, 1, 3, rng), quality=80, subsampling=1)
    save("jpg_420_2x3.jpg", _smooth(3, 2, 3, rng), quality=80, subsampling=2)
    save("jpg_progressive_420.jpg", _smooth(33, 47, 3, rng), quality=85, subsampling=2, progressive=True)
    save("jpg_progressive_444.jpg", _smooth(18, 22, 3, rng), quality=97, subsampling=0, progressive=True,
         optimize=True)
    save("jpg_progressive_grey.jpg", _smooth(25, 19, 1, rng)[..., 0], quality=70, progressive=True)
    save("jpg_grey.jpg", _smooth(17, 30, 1, rng)[..., 0], quality=88)
    save("jpg_restart_blocks.jpg", _smooth(31, 45, 3, rng), quality=90, subsampling=2, restart_marker_blocks=3)
    save("jpg_restart_rows_progressive.jpg", _smooth(40, 24, 3, rng), quality=90, subsampling=1,
         restart_marker_rows=1, progressive=True)
    save("jpg_16bit_tables.jpg", _smooth(20, 26, 3, rng), qtables=[[300 + i for i in range(64)], [2] * 64])
    save("jpg_optimized_huffman.jpg", _smooth(24, 24, 3, rng), quality=60, optimize=True)
    save("jpg_cmyk.jpg", _smooth(14, 21, 3, rng), mode="CMYK", quality=90)
    save("jpg_rgb_ids.jpg", _smooth(16, 20, 3, rng), quality=90, keep_rgb=True)
    return cases


class _BitWriter:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, code, length):
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((code >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)
                self.acc = 0
                self.n = 0

    def flush(self):
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)


def _category(v):
    return 0 if v == 0 else int(abs(v)).bit_length()


def _bits_of(v, s):
    return v if v >= 0 else v + (1 << s) - 1


_ZZ = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14,
       21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60,
       61, 54, 47, 55, 62, 63]


def encode_baseline(planes_full, factors, q=8, restart=0):
    """Minimal baseline JPEG encoder: YCbCr (or grey) planes of shape (H, W),
    per-component sampling factors (h, v), a flat quantiser q, and one 4-bit
    DC code and one 8-bit AC code per symbol (a valid canonical table)."""
    H, W = planes_full[0].shape
    hmax = max(f[0] for f in factors)
    vmax = max(f[1] for f in factors)
    mcux = -(-W // (8 * hmax))
    mcuy = -(-H // (8 * vmax))
    k = np.arange(8)
    C = np.sqrt(2 / 8) * np.cos((2 * k[None, :] + 1) * k[:, None] * np.pi / 16)
    C[0] /= np.sqrt(2)
    comps = []
    for p, (hh, vv) in zip(planes_full, factors):
        hs, vs = hmax // hh, vmax // vv
        cw, chh = mcux * hh * 8, mcuy * vv * 8
        x = np.pad(p.astype(np.float64), ((0, vmax * 8 * mcuy - H), (0, hmax * 8 * mcux - W)), mode="edge")
        sub = x.reshape(x.shape[0] // vs, vs, x.shape[1] // hs, hs).mean(axis=(1, 3))[:chh, :cw]
        comps.append(sub - 128.0)
    # Quantised coefficients per component block, zig-zag order.
    blocks = []
    for c in comps:
        nby, nbx = c.shape[0] // 8, c.shape[1] // 8
        b = c.reshape(nby, 8, nbx, 8).transpose(0, 2, 1, 3)
        d = np.einsum("ij,abjk,lk->abil", C, b, C)
        blocks.append(np.round(d / q).astype(np.int64).reshape(nby, nbx, 64)[:, :, _ZZ])
    ac_syms = set()
    for b in blocks:
        for blk in b.reshape(-1, 64):
            run = 0
            for v in blk[1:]:
                if v == 0:
                    run += 1
                    continue
                while run > 15:
                    ac_syms.add(0xF0)
                    run -= 16
                ac_syms.add(run << 4 | _category(v))
                run = 0
            if run:
                ac_syms.add(0x00)
    ac_syms = sorted(ac_syms | {0x00})
    ac_code = {s: i for i, s in enumerate(ac_syms)}
    dc_code = {s: s for s in range(12)}
    bw = _BitWriter()
    pred = [0] * len(comps)
    n_mcu = 0
    seg = bytearray()
    rst = 0
    # A one-component scan is non-interleaved: its units are single blocks in
    # raster order over the blocks holding samples; otherwise MCUs.
    if len(comps) == 1:
        units = [[(0, by, bx)] for by in range(-(-H // 8)) for bx in range(-(-W // 8))]
    else:
        units = [[(ci, my * vv + yy, mx * hh + xx) for ci, (hh, vv) in enumerate(factors)
                  for yy in range(vv) for xx in range(hh)] for my in range(mcuy) for mx in range(mcux)]
    for unit in units:
        for ci, by, bx in unit:
            blk = blocks[ci][by, bx]
            diff = int(blk[0]) - pred[ci]
            pred[ci] = int(blk[0])
            s = _category(diff)
            bw.put(dc_code[s], 4)
            if s:
                bw.put(_bits_of(diff, s), s)
            run = 0
            for v in blk[1:]:
                v = int(v)
                if v == 0:
                    run += 1
                    continue
                while run > 15:
                    bw.put(ac_code[0xF0], 8)
                    run -= 16
                s = _category(v)
                bw.put(ac_code[run << 4 | s], 8)
                bw.put(_bits_of(v, s), s)
                run = 0
            if run:
                bw.put(ac_code[0x00], 8)
        n_mcu += 1
        if restart and n_mcu % restart == 0 and n_mcu < len(units):
            bw.flush()
            seg += bw.out + bytes([0xFF, 0xD0 + rst])
            rst = (rst + 1) & 7
            bw = _BitWriter()
            pred = [0] * len(comps)
    bw.flush()
    seg += bw.out

    def marker(m, payload):
        return bytes([0xFF, m]) + struct.pack(">H", len(payload) + 2) + payload

    out = b"\xff\xd8" + marker(0xE0, b"JFIF\x00\x01\x01\x00\x00\x01\x00\x01\x00\x00")
    out += marker(0xDB, b"\x00" + bytes([q] * 64))
    sof = struct.pack(">BHHB", 8, H, W, len(comps))
    for i, (hh, vv) in enumerate(factors):
        sof += bytes([i + 1, hh << 4 | vv, 0])
    out += marker(0xC0, sof)
    dc_counts = [0] * 16
    dc_counts[3] = 12
    out += marker(0xC4, bytes([0x00] + dc_counts + list(range(12))))
    ac_counts = [0] * 16
    ac_counts[7] = len(ac_syms)
    out += marker(0xC4, bytes([0x10] + ac_counts + ac_syms))
    if restart:
        out += marker(0xDD, struct.pack(">H", restart))
    sos = bytes([len(comps)])
    for i in range(len(comps)):
        sos += bytes([i + 1, 0x00])
    out += marker(0xDA, sos + b"\x00\x3f\x00") + bytes(seg) + b"\xff\xd9"
    return out


def own_jpeg_cases(rng):
    cases = []
    for name, f, (h, w), restart in (
        ("jpg_own_1x2.jpg", [(1, 2), (1, 1), (1, 1)], (21, 26), 0),
        ("jpg_own_4x1.jpg", [(4, 1), (1, 1), (1, 1)], (13, 45), 0),
        ("jpg_own_3x1.jpg", [(3, 1), (1, 1), (1, 1)], (10, 29), 2),
        ("jpg_own_2x4.jpg", [(2, 4), (1, 1), (2, 2)], (37, 19), 0),
        ("jpg_own_4x4.jpg", [(4, 4), (2, 2), (1, 1)], (40, 33), 1),
        ("jpg_own_grey_2x2.jpg", [(2, 2)], (15, 15), 0),
    ):
        planes = [_smooth(h, w, 1, rng)[..., 0] for _ in f]
        cases.append((name, encode_baseline(planes, f, q=int(rng.integers(2, 9)), restart=restart)))
    return cases


# ------------------------------------------------------------------ PNM
def pnm_cases(rng):
    a = rng.integers(0, 256, (6, 7, 3)).astype(np.uint8)
    g = rng.integers(0, 101, (5, 9)).astype(np.uint8)
    g16 = rng.integers(0, 65536, (4, 3)).astype(">u2")
    return [
        ("pnm_rgb.ppm", b"P6\n# fixture\n7 6\n255\n" + a.tobytes()),
        ("pnm_grey_maxval100.pgm", b"P5 9 5 100\n" + g.tobytes()),
        ("pnm_grey16.pgm", b"P5\n3\n4\n65535\n" + g16.tobytes()),
    ]


def main() -> int:
    if not os.path.isdir("/root/reference"):
        print("the reference tree is not present; fixtures are regenerated only where it is")
        return 1
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    os.makedirs(OUT, exist_ok=True)
    for f in os.listdir(OUT):
        os.remove(os.path.join(OUT, f))
    rng = np.random.default_rng(20240611)
    files = []
    for name, kw in png_cases(rng):
        write_png(os.path.join(OUT, name), rng=rng, **kw)
        files.append(name)
    for name, data in pillow_jpeg_cases(rng) + own_jpeg_cases(rng) + pnm_cases(rng):
        with open(os.path.join(OUT, name), "wb") as f:
            f.write(data)
        files.append(name)
    expected = {}
    with tempfile.TemporaryDirectory() as td:
        for name in files:
            dst = os.path.join(td, "out.bin")
            subprocess.run([STB, os.path.join(OUT, name), dst], check=True)
            raw = open(dst, "rb").read()
            w, h = struct.unpack("<ii", raw[:8])
            expected[name] = np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4)
    np.savez_compressed(os.path.join(OUT, "expected.npz"), **expected)
    print(f"{len(files)} image fixtures written to {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
