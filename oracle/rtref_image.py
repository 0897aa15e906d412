"""Texture-file reader of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Restates what the reference's texture loader returns (Texture::Load,
Src/Core/texture.cu:207-214: stb_image 2.30's stbi_load with four requested
channels) for PNG and binary PNM files, in numpy: RGBA8, row 0 = top.
 - PNG (ISO/IEC 15948): all colour types and bit depths, Adam7, the five row
   filters; 16-bit samples keep their high byte; 1/2/4-bit grey is scaled by
   255 / (2^depth - 1); a tRNS key (compared at 16 bits, or after the grey
   scaling at 8 bits and less) clears alpha; palette alpha from tRNS.
 - PNM (P5/P6): raw samples, no maxval rescaling; of a 16-bit sample the second
   byte (stbi reads the samples in little-endian host order and keeps the high
   half).
JPEG files are not read here; their decoding is pinned by the expected texels
recorded from the reference's own decoder under tests/golden/images/.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

_ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
_GREY_SCALE = {1: 0xFF, 2: 0x55, 4: 0x11, 8: 1, 16: 1}


def _unfilter(data: bytes, rows: int, stride: int, bpp: int) -> np.ndarray:
    out = np.zeros((rows, stride), np.int64)
    prev = np.zeros(stride, np.int64)
    for y in range(rows):
        ft = data[y * (stride + 1)]
        raw = np.frombuffer(data, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        cur = np.zeros(stride, np.int64)
        if ft == 0:
            cur = raw
        elif ft == 2:
            cur = (raw + prev) & 255
        else:
            for i in range(stride):  # left-dependent filters, byte by byte
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 1:
                    p = a
                elif ft == 3:
                    p = (a + b) >> 1
                elif ft == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                else:
                    raise ValueError("PNG: invalid filter type")
                cur[i] = (raw[i] + p) & 255
        out[y] = cur
        prev = cur
    return out


def _samples(rows: np.ndarray, width: int, channels: int, depth: int) -> np.ndarray:
    """Unpack filtered-off rows into (rows, width, channels) sample values."""
    r = rows.shape[0]
    if depth == 16:
        v = (rows[:, 0::2] << 8) | rows[:, 1::2]
        return v[:, :width * channels].reshape(r, width, channels)
    if depth == 8:
        return rows[:, :width * channels].reshape(r, width, channels)
    per = 8 // depth
    shifts = np.array([8 - depth * (k + 1) for k in range(per)], np.int64)
    bits = (rows[:, :, None] >> shifts[None, None, :]) & ((1 << depth) - 1)
    return bits.reshape(r, -1)[:, :width * channels].reshape(r, width, channels)


def decode_png(data: bytes) -> np.ndarray:
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("PNG: bad signature")
    pos = 8
    w = h = depth = color = interlace = None
    palette = None
    trns = None
    idat = b""
    while True:
        length, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if tag == b"IHDR":
            w, h, depth, color, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif tag == b"PLTE":
            palette = np.concatenate([np.frombuffer(body, np.uint8).reshape(-1, 3).astype(np.int64),
                                      np.full((len(body) // 3, 1), 255, np.int64)], axis=1)
        elif tag == b"tRNS":
            trns = body
        elif tag == b"IDAT":
            idat += body
        elif tag == b"IEND":
            break
    ch = _CHANNELS[color]
    raw = zlib.decompress(idat)
    img = np.zeros((h, w, ch), np.int64)
    off = 0
    passes = _ADAM7 if interlace else [(0, 0, 1, 1)]
    bpp = max(1, ch * depth // 8)
    for x0, y0, dx, dy in passes:
        pw = (w - x0 + dx - 1) // dx if w > x0 else 0
        ph = (h - y0 + dy - 1) // dy if h > y0 else 0
        if pw == 0 or ph == 0:
            continue
        stride = (pw * ch * depth + 7) // 8
        rows = _unfilter(raw[off:off + ph * (stride + 1)], ph, stride, bpp)
        off += ph * (stride + 1)
        img[y0::dy, x0::dx] = _samples(rows, pw, ch, depth)
    out = np.zeros((h, w, 4), np.int64)
    out[..., 3] = 255
    if color == 3:
        pal = palette.copy()
        if trns is not None:
            pal[:len(trns), 3] = np.frombuffer(trns, np.uint8)
        out[:] = pal[img[..., 0]]
        return out.astype(np.uint8)
    key = None
    if trns is not None:
        key = np.array(struct.unpack(">" + "H" * ch, trns), np.int64)
    if depth == 16:
        clear = np.all(img == key, axis=2) if key is not None else None
        img = img >> 8
    else:
        img = img * _GREY_SCALE[depth] if color == 0 else img
        clear = np.all(img == ((key & 255) * _GREY_SCALE[depth]) & 255, axis=2) if key is not None else None
    if color in (0, 4):
        out[..., 0] = out[..., 1] = out[..., 2] = img[..., 0]
        if color == 4:
            out[..., 3] = img[..., 1]
    else:
        out[..., :ch] = img
    if clear is not None:
        out[..., 3] = np.where(clear, 0, 255)
    return out.astype(np.uint8)


def decode_pnm(data: bytes) -> np.ndarray:
    ch = 1 if data[1:2] == b"5" else 3
    pos, vals = 2, []
    while len(vals) < 3:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            while data[pos:pos + 1] not in (b"\n", b"\r"):
                pos += 1
            continue
        start = pos
        while data[pos:pos + 1].isdigit():
            pos += 1
        vals.append(int(data[start:pos]))
    w, h, maxval = vals
    pos += 1
    nbytes = 2 if maxval > 255 else 1
    s = np.frombuffer(data, np.uint8, w * h * ch * nbytes, pos).reshape(h, w, ch, nbytes)[..., nbytes - 1]
    out = np.full((h, w, 4), 255, np.uint8)
    out[..., :3] = s if ch == 3 else s[..., :1]
    return out


def read_texture(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"\x89PNG":
        return decode_png(data)
    if data[:2] in (b"P5", b"P6"):
        return decode_pnm(data)
    raise ValueError(f"the oracle reads PNG and PNM textures only: {path}")
