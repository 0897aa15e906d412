"""Randomised texture decoding against the reference's own decoder: PNG and
JPEG files of random size, colour type, bit depth, interlacing, row filters,
quality, subsampling, progressive / optimised / restart options and sampling
factors, decoded by the library and by stb_image 2.30 built from the
reference tree into oracle/_ref/ (`make -C oracle ref`; skipped where that
tree is absent).  The committed fixtures under tests/golden/images/ pin the
same decoders where stb cannot be built."""
import io
import json
import os
import struct
import subprocess

import numpy as np
import pytest

import cudaraytracer_amd as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STB = os.path.join(ROOT, "oracle", "_ref", "stb_decode")

pytestmark = pytest.mark.skipif(not os.path.exists(STB), reason="the reference's stb_image build is not present")


def _fixture_module():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_image_fixtures",
                                                  os.path.join(ROOT, "tests", "golden", "make_image_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _random_files(tmp, n, seed):
    from PIL import Image

    mf = _fixture_module()
    rng = np.random.default_rng(seed)
    files = []
    for i in range(n):
        h, w = int(rng.integers(1, 70)), int(rng.integers(1, 70))
        kind = i % 3
        p = os.path.join(tmp, f"{i}.img")
        if kind == 0:
            ch = 1 if rng.random() < 0.2 else 3
            arr = mf._smooth(h, w, ch, rng)
            kw = dict(quality=int(rng.integers(1, 101)))
            if ch == 3:
                kw["subsampling"] = int(rng.integers(0, 3))
            if rng.random() < 0.5:
                kw["progressive"] = True
            if rng.random() < 0.3:
                kw["optimize"] = True
            if rng.random() < 0.3:
                kw["restart_marker_blocks"] = int(rng.integers(1, 6))
            b = io.BytesIO()
            Image.fromarray(arr[..., 0] if ch == 1 else arr).save(b, "JPEG", **kw)
            open(p, "wb").write(b.getvalue())
        elif kind == 1:
            layouts = [[(2, 2), (1, 1), (1, 1)], [(1, 2), (1, 1), (1, 1)], [(4, 1), (1, 1), (1, 1)],
                       [(4, 2), (2, 1), (1, 1)], [(3, 3), (1, 1), (3, 1)], [(2, 2)], [(4, 4)]]
            f = layouts[int(rng.integers(0, len(layouts)))]
            planes = [mf._smooth(h, w, 1, rng)[..., 0] for _ in f]
            open(p, "wb").write(mf.encode_baseline(planes, f, q=int(rng.integers(1, 12)),
                                                   restart=int(rng.integers(0, 4))))
        else:
            color = int(rng.choice([0, 2, 3, 4, 6]))
            depth = int(rng.choice({0: [1, 2, 4, 8, 16], 2: [8, 16], 3: [1, 2, 4, 8], 4: [8, 16], 6: [8, 16]}[color]))
            ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[color]
            kw = dict(color=color, depth=depth, interlace=bool(rng.random() < 0.4), split=int(rng.integers(1, 4)))
            if color == 3:
                n_pal = int(rng.integers(1, (1 << depth) + 1))
                kw["palette"] = rng.integers(0, 256, (n_pal, 3))
                kw["samples"] = rng.integers(0, n_pal, (h, w, 1))
                if rng.random() < 0.5:
                    kw["trns"] = bytes(rng.integers(0, 256, int(rng.integers(1, n_pal + 1))).astype(np.uint8))
            else:
                s = rng.integers(0, 1 << depth, (h, w, ch))
                kw["samples"] = s
                if color in (0, 2) and rng.random() < 0.5:
                    key = s[rng.integers(0, h), rng.integers(0, w)]
                    kw["trns"] = struct.pack(">" + "H" * ch, *[int(v) for v in key])
            mf.write_png(p, rng=rng, **kw)
        files.append(p)
    return files


def test_random_images_match_reference_decoder(tmp_path):
    files = _random_files(str(tmp_path), 150, 77)
    with open(tmp_path / "quad.obj", "w") as f:
        f.write("v -1 0 -1\nv 1 0 -1\nv 1 0 1\nf 1 3 2\n")
    doc = {"camera": {"type": "pinhole", "resolution": [8, 8], "fov": 40, "spp": 1, "max_bounces": 2, "rr": 0.8,
                      "direction": [0, -1, 0]},
           "textures": [{"name": str(i), "path": p} for i, p in enumerate(files)],
           "materials": [{"name": "m", "type": "diffuse", "basecolor": [0.5, 0.5, 0.5]}],
           "objects": [{"name": "quad", "path": "quad.obj", "material": "m"}], "lights": []}
    path = str(tmp_path / "s.json")
    with open(path, "w") as f:
        json.dump(doc, f)
    scene = rt.Scene(path)
    for i, p in enumerate(files):
        out = p + ".stb"
        subprocess.run([STB, p, out], check=True)
        raw = open(out, "rb").read()
        w, h = struct.unpack("<ii", raw[:8])
        want = np.frombuffer(raw[8:], np.uint8).reshape(h, w, 4)
        got = scene.texture(i)
        assert got.shape == want.shape and np.array_equal(got, want), f"file {i}"
