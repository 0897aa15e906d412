"""Write the Radiance RGBE (.hdr) fixtures under tests/golden/hdr/.

Inputs: small .hdr files covering what the reference's loader distinguishes:
new-style run-length scanlines (runs and literal dumps of every length), the
"#?RADIANCE" and "#?RGBE" signatures, extra and over-long header lines, the
spacing the resolution line allows, widths below 8 (always flat), a first
scanline that is not run-length encoded (the whole file is then read flat), a
flat scanline after run-length ones (flat decoding restarts at row 0, pixel 1),
a length byte with its high bit set, and every exponent edge (0, denormal
results, 255).  Files the loader refuses are kept too, with its reason.

Expected outputs: what the reference's HDRTexture::Load receives
(Src/Core/texture.cu:17-25: stbi_loadf(path, &w, &h, &c, 4), float RGBA,
row 0 = first scanline), from the reference's vendored stb_image 2.30 built
from the reference tree (`make -C oracle ref`, oracle/stb_decode.c -f).

Run where the reference tree is present:
    python tests/golden/make_hdr_fixtures.py
It rewrites hdr/*.hdr, hdr/expected.npz (one float32 array of shape (h, w, 4)
per decodable file) and hdr/refused.json (file name -> stb's failure reason).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "hdr")
STB = os.path.join(ROOT, "oracle", "_ref", "stb_decode")


def header(w: int, h: int, sig: bytes = b"#?RADIANCE", extra: list[bytes] | None = None,
           res: bytes | None = None) -> bytes:
    lines = [sig, b"# written by make_hdr_fixtures.py"] + (extra or []) + [b"FORMAT=32-bit_rle_rgbe", b""]
    return b"\n".join(lines) + b"\n" + (res if res is not None else b"-Y %d +X %d" % (h, w)) + b"\n"


def encode_channel(vals: np.ndarray) -> bytes:
    """One channel of a new-style scanline: runs (count + 128, value) for
    repeats of 3 or more (at most 127), literal dumps (count, bytes) of at most
    128 otherwise."""
    out = bytearray()
    i, n = 0, len(vals)
    lit = []

    def flush():
        while lit:
            chunk = lit[:128]
            del lit[:128]
            out.append(len(chunk))
            out.extend(chunk)

    while i < n:
        j = i
        while j < n and vals[j] == vals[i] and j - i < 127:
            j += 1
        if j - i >= 3:
            flush()
            out.append(128 + (j - i))
            out.append(int(vals[i]))
            i = j
        else:
            lit.append(int(vals[i]))
            i += 1
    flush()
    return bytes(out)


def rle_scanline(px: np.ndarray) -> bytes:
    w = px.shape[0]
    return bytes([2, 2, w >> 8, w & 255]) + b"".join(encode_channel(px[:, c]) for c in range(4))


def rle_image(px: np.ndarray) -> bytes:
    return b"".join(rle_scanline(px[y]) for y in range(px.shape[0]))


def rgbe_from_float(rgb: np.ndarray) -> np.ndarray:
    """Ward's RGBE encoding of float RGB (h, w, 3) -> uint8 (h, w, 4)."""
    v = rgb.max(axis=2).astype(np.float64)
    m, e = np.frexp(v)
    out = np.zeros(rgb.shape[:2] + (4,), np.uint8)
    ok = v > 1e-32
    scale = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    out[..., :3] = np.clip(np.floor(rgb * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def random_pixels(rng, h, w, runs=True):
    px = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    px[..., 3] = rng.integers(100, 160, (h, w))
    if runs:  # stretches of repeated bytes so the encoder writes runs
        for y in range(h):
            a = int(rng.integers(0, max(1, w - 4)))
            b = min(w, a + int(rng.integers(3, 40)))
            px[y, a:b, int(rng.integers(0, 4))] = int(rng.integers(0, 256))
    zero = rng.random((h, w)) < 0.1
    px[zero, 3] = 0
    return px


def cases() -> dict[str, bytes]:
    rng = np.random.default_rng(2024)
    c: dict[str, bytes] = {}
    px = random_pixels(rng, 5, 16)
    c["rle_runs_16x5.hdr"] = header(16, 5, extra=[b"EXPOSURE=1.0", b"GAMMA=1"]) + rle_image(px)
    px = random_pixels(rng, 3, 9)
    c["rgbe_signature_9x3.hdr"] = header(9, 3, sig=b"#?RGBE") + rle_image(px)
    # literal dumps of 128 and runs of 127 in one long scanline
    px = random_pixels(rng, 2, 300, runs=False)
    px[0, 10:137, 1] = 7
    px[1, :, 2] = 200
    c["rle_long_300x2.hdr"] = header(300, 2) + rle_image(px)
    px = random_pixels(rng, 4, 5, runs=False)
    c["flat_narrow_5x4.hdr"] = header(5, 4) + px.tobytes()
    px = random_pixels(rng, 3, 1, runs=False)
    c["flat_narrow_1x3.hdr"] = header(1, 3) + px.tobytes()
    px = random_pixels(rng, 3, 12, runs=False)
    px[0, 0, 0] = 200  # first scanline is not run-length encoded: the whole file is flat
    c["flat_first_scanline_12x3.hdr"] = header(12, 3) + px.tobytes()
    px = random_pixels(rng, 4, 10)
    flat = random_pixels(rng, 4, 10, runs=False)
    flat[0, 0, 0] = 77  # row 2's header is a pixel: flat decoding restarts at row 0, pixel 1
    c["rle_then_flat_restart_10x4.hdr"] = header(10, 4) + rle_image(px[:2]) + flat.tobytes()
    px = random_pixels(rng, 2, 8, runs=False)
    px[0, 0, :3] = [2, 2, 0x88]  # length byte with the high bit set: not run-length encoded
    c["rle_length_highbit_8x2.hdr"] = header(8, 2) + px.tobytes()
    ex = np.zeros((4, 8, 4), np.uint8)
    ex[..., 0] = [[255, 1, 128, 0, 17, 255, 3, 64]] * 4
    ex[..., 1] = [[0, 255, 128, 5, 19, 254, 2, 65]] * 4
    ex[..., 2] = [[1, 2, 3, 255, 23, 253, 1, 66]] * 4
    ex[..., 3] = [[0, 1, 2, 8, 127, 128, 136, 255], [9, 10, 11, 126, 129, 135, 137, 254],
                  [1, 1, 1, 1, 0, 0, 255, 255], [200, 60, 20, 3, 2, 1, 0, 250]]
    c["exponents_8x4.hdr"] = header(8, 4) + rle_image(ex)
    px = random_pixels(rng, 2, 9)
    c["resolution_spacing_9x2.hdr"] = header(9, 2, res=b"-Y   2   +X  9") + rle_image(px)
    px = random_pixels(rng, 2, 8)
    c["long_header_line_8x2.hdr"] = header(8, 2, extra=[b"#" + b"x" * 1500]) + rle_image(px)
    y, x = np.mgrid[0:32, 0:64].astype(np.float32)
    sky = np.stack([0.2 + 0.6 * y / 32, 0.4 + 0.4 * y / 32, 1.0 - 0.3 * y / 32], -1)
    sky += 1e4 * np.exp(-((x - 40) ** 2 + (y - 8) ** 2) / 2.0)[..., None] * np.array([1.0, 0.9, 0.7], np.float32)
    c["sky_sun_64x32.hdr"] = header(64, 32) + rle_image(rgbe_from_float(sky))
    # refused by the loader
    px = random_pixels(rng, 2, 8)
    good = rle_image(px)
    c["bad_no_format_8x2.hdr"] = b"#?RADIANCE\n# no format line\n\n-Y 2 +X 8\n" + good
    c["bad_layout_plus_y_8x2.hdr"] = header(8, 2, res=b"+Y 2 +X 8") + good
    c["bad_signature_crlf_8x2.hdr"] = header(8, 2).replace(b"#?RADIANCE\n", b"#?RADIANCE\r\n") + good
    c["bad_scanline_length_8x2.hdr"] = header(8, 2) + bytes([2, 2, 0, 9]) + good[4:]
    c["bad_run_overflow_8x2.hdr"] = header(8, 2) + bytes([2, 2, 0, 8, 128 + 9, 1]) + good[6:]
    c["bad_zero_dump_8x2.hdr"] = header(8, 2) + bytes([2, 2, 0, 8, 0]) + good[5:]
    c["bad_negative_height_8x2.hdr"] = header(8, 2, res=b"-Y -2 +X 8") + good
    return c


def stb_decode(path: str):
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "o.bin")
        r = subprocess.run([STB, "-f", path, out], capture_output=True, text=True)
        if r.returncode == 2:
            return None, r.stderr.strip()
        if r.returncode != 0:
            raise RuntimeError(f"stb_decode failed on {path}: {r.stderr}")
        raw = open(out, "rb").read()
    w, h = np.frombuffer(raw[:8], np.int32)
    return np.frombuffer(raw[8:], np.float32).reshape(h, w, 4).copy(), None


def main() -> int:
    if not os.path.exists(STB):
        print(f"{STB} is missing: run `make -C oracle ref` where the reference tree is present", file=sys.stderr)
        return 1
    os.makedirs(OUT, exist_ok=True)
    expected, refused = {}, {}
    for name, data in cases().items():
        path = os.path.join(OUT, name)
        with open(path, "wb") as f:
            f.write(data)
        arr, reason = stb_decode(path)
        if arr is None:
            refused[name] = reason
        else:
            expected[name] = arr
    if any(n.startswith("bad_") for n in expected) or any(not n.startswith("bad_") for n in refused):
        raise RuntimeError(f"unexpected decode outcome: refused {sorted(refused)}")
    np.savez_compressed(os.path.join(OUT, "expected.npz"), **expected)
    with open(os.path.join(OUT, "refused.json"), "w") as f:
        json.dump(refused, f, indent=1, sort_keys=True)
    print(f"{len(expected)} decodable and {len(refused)} refused files written to {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
