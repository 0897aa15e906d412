"""BVH refit on the CPU: the oracle's refit (ref_scene_refit) pinned by what the
mathematics fixes, and the library's host refit (rt_scene_refit on a scene that
is not resident on a GPU) against the oracle.

The reference has no refit (it reconstructs the Scene, Src/Core/scene.cpp:50-57,
and rebuilds the tree), so its semantics are this library's: the same tree,
a leaf's box the union of its triangles' boxes (Bounds3D(v0, v1, v2),
Src/Core/bounds.h:22-26), an interior node's box the union of its children's.
Pins: (1) refit to the loaded positions gives back the build's boxes (the build
forms each node's box from its triangles, bvh.cu:61-64); (2) every box equals
the min/max over its subtree's vertices, computed by brute force; (3) boxes
commute with a monotone map of the coordinates (x -> fl(fl(x*s) + d), s > 0);
(4) after a refit the oracle finds the same closest hits as a fresh build of
the moved scene."""
import numpy as np
import pytest

import cudaraytracer_amd as rt

F32 = np.float32


def _boxes(nodes):
    """(n, 6) float32 lo.xyz, hi.xyz of reference-layout nodes (8 words each)."""
    return np.ascontiguousarray(nodes[:, 2:8]).view(np.float32)


def _ranges(nodes):
    """Ordered-triangle range [start, end) of every node's subtree (DFS pre-order)."""
    n = nodes.shape[0]
    start = np.zeros(n, np.int64)
    end = np.zeros(n, np.int64)
    for i in range(n - 1, -1, -1):
        if nodes[i, 0] >= 0:
            start[i], end[i] = nodes[i, 0], nodes[i, 0] + nodes[i, 1]
        else:
            start[i], end[i] = start[i + 1], end[nodes[i, 1]]
    return start, end


def _brute_boxes(nodes, order, pos):
    """min / max over each node's subtree vertices, by numpy reductions."""
    start, end = _ranges(nodes)
    v = pos[order].reshape(-1, 3, 3)
    lo = v.min(axis=1)
    hi = v.max(axis=1)
    out = np.zeros((nodes.shape[0], 6), np.float32)
    for i in range(nodes.shape[0]):
        out[i, :3] = lo[start[i]:end[i]].min(axis=0)
        out[i, 3:] = hi[start[i]:end[i]].max(axis=0)
    return out


def deform(pos, mesh, seed=0, jitter=0.01, scale_mesh=None):
    """Moved vertices for tests: a random translation per mesh plus a small jitter
    per vertex, and one mesh (default: the last triangle's) scaled by 1.25 about
    its centre."""
    rng = np.random.default_rng(seed)
    p = pos.reshape(-1, 3, 3).astype(np.float64).copy()
    for m in np.unique(mesh):
        sel = mesh == m
        p[sel] += rng.uniform(-0.05, 0.05, 3)
    p += rng.uniform(-jitter, jitter, p.shape)
    last = mesh == (mesh[-1] if scale_mesh is None else scale_mesh)
    c = p[last].reshape(-1, 3).mean(axis=0)
    p[last] = c + 1.25 * (p[last] - c)
    return p.reshape(-1, 9).astype(np.float32)


@pytest.fixture(scope="module")
def spheres(scene_dir, oracle_lib):
    a = oracle_lib.parse_scene(scene_dir["cornell_spheres_128"])
    return a


def test_oracle_refit_to_loaded_positions_gives_the_build(spheres, oracle_lib):
    rs = oracle_lib.RefScene(spheres)
    n0, o0 = rs.bvh()
    rs.refit(spheres.pos)
    n1, o1 = rs.bvh()
    assert np.array_equal(o0, o1) and np.array_equal(n0[:, :2], n1[:, :2])
    assert np.array_equal(_boxes(n0), _boxes(n1))  # as values: -0 == +0


def test_oracle_refit_boxes_are_the_subtree_extents(spheres, oracle_lib):
    rs = oracle_lib.RefScene(spheres)
    moved = deform(spheres.pos, spheres.mesh, seed=1)
    rs.refit(moved)
    nodes, order = rs.bvh()
    assert np.array_equal(_boxes(nodes), _brute_boxes(nodes, order, moved))
    # ... and the build's boxes of the loaded positions are not those extents any more.
    assert not np.array_equal(_boxes(nodes), _brute_boxes(nodes, order, spheres.pos.reshape(-1, 9)))


@pytest.mark.parametrize("s,d", [(F32(2.0), F32(0.5)), (F32(0.75), F32(-3.25)), (F32(1.0009765625), F32(0.0))])
def test_oracle_refit_commutes_with_monotone_maps(spheres, oracle_lib, s, d):
    """Rounding is monotone, so min/max commute with x -> fl(fl(x*s)+d) for s > 0."""
    rs = oracle_lib.RefScene(spheres)
    n0, _ = rs.bvh()
    rs.refit((spheres.pos.reshape(-1, 9) * s + d).astype(np.float32))
    n1, _ = rs.bvh()
    assert np.array_equal(_boxes(n1), (_boxes(n0) * s + d).astype(np.float32))


def test_oracle_refit_finds_the_fresh_builds_hits(spheres, oracle_lib):
    import dataclasses

    moved = deform(spheres.pos, spheres.mesh, seed=2)
    rs = oracle_lib.RefScene(spheres)
    rs.refit(moved)
    fresh = oracle_lib.RefScene(dataclasses.replace(spheres, pos=moved.reshape(spheres.pos.shape)))
    rng = np.random.default_rng(4)
    n = 4000
    lo, hi = moved.reshape(-1, 3).min(0), moved.reshape(-1, 3).max(0)
    o = rng.uniform(lo, hi, (n, 3))
    dvec = rng.standard_normal((n, 3))
    dvec /= np.linalg.norm(dvec, axis=1, keepdims=True)
    rays = np.concatenate([o, np.zeros((n, 1)), dvec, np.full((n, 1), 1e8)], axis=1).astype(np.float32)
    t1, _, _, p1, _ = rs.trace(rays)
    t2, _, _, p2, _ = fresh.trace(rays)
    assert (p1 >= 0).mean() > 0.3
    assert np.array_equal(p1 >= 0, p2 >= 0)
    assert np.array_equal(t1, t2)
    # Triangle ids differ between the two trees' orders; the scene triangle does not.
    _, o1 = rs.bvh()
    _, o2 = fresh.bvh()
    h = p1 >= 0
    assert np.array_equal(o1[p1[h]], o2[p2[h]])


# --------------------------------------------------------------- library
def test_library_host_refit_matches_oracle(scene_dir, oracle_lib, spheres):
    s = rt.Scene(scene_dir["cornell_spheres_128"], bvh_builder=rt.BVH_HOST)
    pos, nrm, _, mesh = s.triangles()
    moved = deform(pos, mesh, seed=3)
    new_nrm = np.roll(nrm, 1, axis=0)
    s.refit(moved, new_nrm)
    info = s.info()
    assert info.refit_on_gpu == 0 and info.refit_ms > 0
    p2, n2, _, _ = s.triangles()
    assert np.array_equal(p2, moved) and np.array_equal(n2, new_nrm)
    rs = oracle_lib.RefScene(spheres)
    rs.refit(moved, new_nrm)
    rn, ro = rs.bvh()
    gn, go = s.bvh()
    assert np.array_equal(go, ro)
    assert np.array_equal(gn, rn)  # bit for bit, signs of zero included
    fn, _ = s.bvh(from_packed=True)
    assert np.array_equal(fn, gn)
    assert np.array_equal(np.array(info.bounds_lo, np.float32), _boxes(rn)[0, :3])
    assert np.array_equal(np.array(info.bounds_hi, np.float32), _boxes(rn)[0, 3:])


def test_library_refit_scaling_by_two_scales_areas_by_four(scene_dir):
    s = rt.Scene(scene_dir["cornell_spheres_128"], bvh_builder=rt.BVH_HOST)
    pos, _, _, _ = s.triangles()
    _, area0 = s.meshes()
    lo0, hi0 = np.array(s.info().bounds_lo, np.float32), np.array(s.info().bounds_hi, np.float32)
    s.refit(pos * F32(2))
    _, area1 = s.meshes()
    assert np.array_equal(area1, area0 * F32(4))  # powers of two: every step is exact
    assert np.array_equal(np.array(s.info().bounds_lo, np.float32), lo0 * F32(2))
    assert np.array_equal(np.array(s.info().bounds_hi, np.float32), hi0 * F32(2))


def test_library_refit_to_loaded_positions_keeps_the_tree(scene_dir):
    s = rt.Scene(scene_dir["cornell_spheres_128"], bvh_builder=rt.BVH_HOST)
    n0, o0 = s.bvh()
    pos, _, _, _ = s.triangles()
    s.refit(pos)
    n1, o1 = s.bvh()
    assert np.array_equal(o0, o1) and np.array_equal(n0[:, :2], n1[:, :2])
    assert np.array_equal(_boxes(n0), _boxes(n1))


def test_library_refit_rejects_bad_input_and_keeps_the_scene(scene_dir):
    s = rt.Scene(scene_dir["cornell_small"], bvh_builder=rt.BVH_HOST)
    pos, nrm, _, _ = s.triangles()
    n0, _ = s.bvh()
    with pytest.raises(rt.RtError) as e:
        s.refit(pos[:-1])
    assert e.value.code == -1 and "triangles given" in str(e.value)
    bad = pos.copy()
    bad[5, 4] = np.nan
    with pytest.raises(rt.RtError) as e:
        s.refit(bad)
    assert e.value.code == -1 and "non-finite" in str(e.value)
    bad[5, 4] = np.inf
    with pytest.raises(rt.RtError):
        s.refit(bad)
    p2, n2, _, _ = s.triangles()
    assert np.array_equal(p2, pos) and np.array_equal(n2, nrm)
    assert np.array_equal(s.bvh()[0], n0)
    assert s.info().refit_ms == 0
