"""rt_trace_rays_noalloc, the closest-hit entry point that allocates no device
memory (its staging is one fixed device array of the library), against the
allocating rt_trace_rays: identical hits and traversal counts, bit for bit,
on two calls in a row (the second smaller, so stale contents of the shared
array would show), at the size limit, and its refusals (more rays than the
array holds, a scene not yet resident on the device)."""
import ctypes

import numpy as np
import pytest

import cudaraytracer_amd as rt
from cudaraytracer_amd import scenes

MAX_RAYS = 262144  # RT_NOALLOC_MAX_RAYS
RT_ERR_INVALID_ARG = -1


def _rays(n, seed, lo, hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.uniform(lo, hi, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 0.0, d, 1e8
    return rays


def _noalloc(scene, rays):
    n = rays.shape[0]
    hits = np.full((n, 4), np.nan, np.float32)
    counts = np.full((n, 2), 0xFFFFFFFF, np.uint32)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = rt.lib().rt_trace_rays_noalloc(scene._h, np.ascontiguousarray(rays).ctypes.data_as(fp), n,
                                        hits.ctypes.data_as(fp),
                                        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return rc, hits, counts


@pytest.mark.gpu
def test_noalloc_trace_matches_allocating_entry_point(tmp_path):
    path = scenes.write_scene(scenes.cornell_spheres(width=32, height=32, spp=1), str(tmp_path), "spheres")
    scene = rt.Scene(path)
    info = scene.info()
    lo, hi = float(min(info.bounds_lo)), float(max(info.bounds_hi))
    first = _rays(60000, 1, lo, hi)
    # not resident on the device yet: refused before any device work
    rc, _, _ = _noalloc(scene, first[:16])
    assert rc == RT_ERR_INVALID_ARG and b"resident" in rt.lib().rt_last_error()
    for rays in (first, _rays(777, 2, lo, hi), _rays(MAX_RAYS, 3, lo, hi), _rays(5, 4, lo, hi)):
        t, b1, b2, prim, ref_counts = scene.trace(rays)  # the allocating entry point (puts the scene on the device)
        ref_hits = np.stack([t, b1, b2, prim.view(np.float32)], axis=1)
        rc, hits, counts = _noalloc(scene, rays)
        assert rc == rt.RT_OK, rt.lib().rt_last_error()
        assert np.array_equal(hits.view(np.int32), ref_hits.view(np.int32))
        assert np.array_equal(counts, ref_counts)
        assert (hits[:, 3].view(np.int32) >= 0).sum() > 0
    # more rays than the array holds: refused with the limit named
    rc, _, _ = _noalloc(scene, np.zeros((MAX_RAYS + 1, 8), np.float32))
    assert rc == RT_ERR_INVALID_ARG and str(MAX_RAYS).encode() in rt.lib().rt_last_error()
    # zero rays is a no-op
    rc, _, _ = _noalloc(scene, np.zeros((0, 8), np.float32))
    assert rc == rt.RT_OK
