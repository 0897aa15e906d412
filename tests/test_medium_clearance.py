"""Clearance grids of random-walk media: every cell value is a lower bound of
the distance from any point of the cell to the medium's mesh (so a walk step
shorter than it can skip its query without changing the walk), and the grid
is not trivially zero (deep cells clear the surface).  Distances are exact
point-triangle distances in float64 over all of the mesh's triangles."""
import numpy as np
import pytest

import cudaraytracer_amd as rt
from cudaraytracer_amd import scenes


def _point_triangle_dist(p, a, b, c):
    """Exact distance from each point p (m, 3) to each triangle (n, 3 vertices): (m, n)."""
    ab, ac = b - a, c - a
    ap = p[:, None, :] - a[None]
    d1 = np.einsum("nk,mnk->mn", ab, ap)
    d2 = np.einsum("nk,mnk->mn", ac, ap)
    # project onto the plane, clamp barycentrics by regions (brute force over edges and face)
    n = np.cross(ab, ac)
    nn = np.einsum("nk,nk->n", n, n)
    dist_plane = np.einsum("nk,mnk->mn", n, ap) / np.sqrt(nn)[None]
    # barycentric of the projection
    aa = np.einsum("nk,nk->n", ab, ab)
    bb = np.einsum("nk,nk->n", ac, ac)
    abac = np.einsum("nk,nk->n", ab, ac)
    den = aa * bb - abac * abac
    v = (bb[None] * d1 - abac[None] * d2) / den[None]
    w = (aa[None] * d2 - abac[None] * d1) / den[None]
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)

    def seg(p0, p1):
        e = p1 - p0
        t = np.clip(np.einsum("nk,mnk->mn", e, p[:, None, :] - p0[None]) / np.einsum("nk,nk->n", e, e)[None], 0, 1)
        q = p0[None] + t[..., None] * e[None]
        return np.linalg.norm(p[:, None, :] - q, axis=-1)

    edge = np.minimum(np.minimum(seg(a, b), seg(b, c)), seg(c, a))
    return np.where(inside, np.abs(dist_plane), edge)


@pytest.mark.parametrize("name", ["cornell_sss_walk", "terrain_small"])
def test_clearance_is_a_lower_bound(scene_dir, name):
    scene = rt.Scene(scene_dir[name], bvh_builder=rt.BVH_HOST)
    pos = scene.triangles()[0].astype(np.float64)
    mesh4, _ = scene.meshes()
    media = 0
    rng = np.random.default_rng(7)
    for m in range(mesh4.shape[0]):
        g = scene.medium(m)
        if g is None:
            continue
        media += 1
        lo, inv, cells = g
        n = cells.shape[0]
        start, count = int(mesh4[m, 2]), int(mesh4[m, 3])
        tri = pos[start:start + count].reshape(-1, 3, 3)
        hi = lo + np.float32(n) / inv
        pts = rng.uniform(lo, hi, (3000, 3)).astype(np.float32)
        cell = np.floor((pts - lo) * inv).astype(np.int64)  # FP32 arithmetic, as the walk does
        ok = np.all((cell >= 0) & (cell < n), axis=1)
        pts, cell = pts[ok], cell[ok]
        bound = cells[cell[:, 2], cell[:, 1], cell[:, 0]].astype(np.float64)
        exact = np.concatenate([_point_triangle_dist(pts[k:k + 200].astype(np.float64), tri[:, 0], tri[:, 1],
                                                     tri[:, 2]).min(1) for k in range(0, len(pts), 200)])
        assert np.all(bound <= exact), float(np.max(bound - exact))
        # deep inside a ball the bound is most of the true distance
        assert cells.max() > 0.25 * exact.max()
    assert media >= 1


def test_media_only_for_random_walk_meshes(scene_dir):
    scene = rt.Scene(scene_dir["terrain_small_burley"], bvh_builder=rt.BVH_HOST)
    mesh4, _ = scene.meshes()
    assert all(scene.medium(m) is None for m in range(mesh4.shape[0]))
    scene = rt.Scene(scene_dir["terrain_small"], bvh_builder=rt.BVH_HOST)
    assert [m for m in range(scene.meshes()[0].shape[0]) if scene.medium(m) is not None] == [1]


def test_clearance_follows_a_refit(tmp_path):
    spec = scenes.cornell_sss_walk(width=8, height=8, spp=1)
    scene = rt.Scene(scenes.write_scene(spec, str(tmp_path), "w"), bvh_builder=rt.BVH_HOST)
    mesh4, _ = scene.meshes()
    m = [k for k in range(mesh4.shape[0]) if scene.medium(k) is not None][0]
    lo0, inv0, c0 = scene.medium(m)
    pos, nrm, _, _ = scene.triangles()
    start, count = int(mesh4[m, 2]), int(mesh4[m, 3])
    p = pos.copy()
    v = p[start:start + count].reshape(-1, 3, 3)
    centre = v.reshape(-1, 3).mean(0)
    p[start:start + count] = ((v - centre) * 1.5 + centre).reshape(-1, 9)  # grow the ball
    scene.refit(p)
    lo1, inv1, c1 = scene.medium(m)
    assert np.all(lo1 < lo0) and c1.max() > 1.3 * c0.max()
