"""The GPU SAH builder against the host builder and the oracle: the node
array (reference layout, DFS pre-order), the triangle order and the packed
64-byte layout must be bit-identical, including fallback splits, signed zeros,
collinear (zero-area) nodes and single-leaf trees."""
import numpy as np
import pytest

import cudaraytracer_amd as rt
from cudaraytracer_amd import scenes

pytestmark = pytest.mark.gpu


def _soup_scene(tmp_path, pos, name):
    spec = scenes.SceneSpec(camera=scenes._camera(8, 8, 1, 2), materials=scenes._base_materials(),
                            meshes=[scenes.MeshSpec("soup", "white", np.asarray(pos, np.float32))], lights=[])
    return scenes.write_scene(spec, str(tmp_path / name), name)


def _assert_same_tree(path, oracle_lib=None, pos=None):
    host = rt.Scene(path, bvh_builder=rt.BVH_HOST)
    gpu = rt.Scene(path, bvh_builder=rt.BVH_GPU)
    hi, gi = host.info(), gpu.info()
    assert gi.bvh_on_gpu == 1 and hi.bvh_on_gpu == 0
    print(f"{path}: {hi.num_triangles} tris, host {hi.bvh_build_ms:.2f} ms, gpu {gi.bvh_build_ms:.2f} ms", flush=True)
    assert (gi.num_bvh_nodes, gi.bvh_max_depth, gi.bvh_fallback_splits) == \
        (hi.num_bvh_nodes, hi.bvh_max_depth, hi.bvh_fallback_splits)
    hn, ho = host.bvh()
    gn, go = gpu.bvh()
    assert np.array_equal(gn, hn), f"{np.count_nonzero((gn != hn).any(axis=1))} nodes differ"
    assert np.array_equal(go, ho)
    gp, _ = gpu.bvh(from_packed=True)
    assert np.array_equal(gp, hn)
    assert list(gi.bounds_lo) == list(hi.bounds_lo) and list(gi.bounds_hi) == list(hi.bounds_hi)
    if oracle_lib is not None and pos is not None:
        rs = oracle_lib.RefScene(oracle_lib.parse_scene(path))
        rn, ro = rs.bvh()
        assert np.array_equal(gn, rn) and np.array_equal(go, ro)
    return hi


@pytest.mark.parametrize("name", ["cornell_small", "cornell_spheres", "terrain_small", "textured_cornell"])
def test_gpu_build_matches_host_on_workloads(scene_dir, name):
    _assert_same_tree(scene_dir[name])


@pytest.mark.parametrize("n,seed", [(1, 1), (4, 2), (5, 3), (37, 4), (4097, 5), (70000, 6), (300000, 7)])
def test_gpu_build_matches_host_and_oracle_on_soups(tmp_path, oracle_lib, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    pos = (c + rng.normal(0, 0.05, (n, 3, 3))).astype(np.float32)
    _assert_same_tree(_soup_scene(tmp_path, pos, f"soup{n}"), oracle_lib, pos)


def test_gpu_build_degenerate_inputs(tmp_path, oracle_lib):
    rng = np.random.default_rng(9)
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    # Identical triangles: all centroids coincide, every split is a fallback.
    info = _assert_same_tree(_soup_scene(tmp_path, np.tile(tri, (300, 1, 1)), "dup"))
    assert info.bvh_fallback_splits > 0
    # Signed zeros on the bounds: the first triangle in range order decides the sign.
    z = rng.uniform(-1, 1, (500, 3, 3)).astype(np.float32)
    z[rng.random((500, 3, 3)) < 0.3] = 0.0
    z[rng.random((500, 3, 3)) < 0.3] = -0.0
    _assert_same_tree(_soup_scene(tmp_path, z, "zeros"))
    # Zero-area triangles on one line (collinear nodes: zero surface area).
    t = rng.uniform(-1, 1, (400, 3, 1)).astype(np.float32)
    line = np.concatenate([t, np.zeros_like(t), np.zeros_like(t)], axis=2)
    _assert_same_tree(_soup_scene(tmp_path, line, "line"))
    # A flat grid of many equal-size triangles (many tied bucket costs).
    g = np.stack(np.meshgrid(np.arange(40), np.arange(40)), -1).reshape(-1, 2).astype(np.float32)
    quad = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    flat = np.concatenate([(g[:, None, :] + quad[None]) , np.zeros((g.shape[0], 3, 1), np.float32)], axis=2)
    _assert_same_tree(_soup_scene(tmp_path, flat, "flat"))


def test_gpu_build_is_the_automatic_choice_for_large_scenes(tmp_path):
    rng = np.random.default_rng(11)
    n = 1 << 16
    pos = (rng.uniform(-1, 1, (n, 1, 3)) + rng.normal(0, 0.01, (n, 3, 3))).astype(np.float32)
    path = _soup_scene(tmp_path, pos, "auto")
    assert rt.Scene(path).info().bvh_on_gpu == 1
    assert rt.Scene(path, bvh_builder=rt.BVH_HOST).info().bvh_on_gpu == 0


def test_gpu_build_beside_a_render(scene_dir, tmp_path):
    """A build runs on its own stream while a render is in flight on another:
    both results stay what they are alone."""
    rng = np.random.default_rng(12)
    n = 100000
    pos = (rng.uniform(-1, 1, (n, 1, 3)) + rng.normal(0, 0.02, (n, 3, 3))).astype(np.float32)
    path = _soup_scene(tmp_path, pos, "beside")
    alone = rt.Renderer(rt.Scene(scene_dir["cornell_spheres_128"]))
    alone.render(4)
    want = alone.read_accum(rt.RT_ACCUM_MEAN)
    r = rt.Renderer(rt.Scene(scene_dir["cornell_spheres_128"]))
    r.render(4)  # asynchronous on the renderer's stream
    built = rt.Scene(path, bvh_builder=rt.BVH_GPU)
    host = rt.Scene(path, bvh_builder=rt.BVH_HOST)
    r.synchronize()
    assert np.array_equal(r.read_accum(rt.RT_ACCUM_MEAN).view(np.int32), want.view(np.int32))
    gn, go = built.bvh()
    hn, ho = host.bvh()
    assert np.array_equal(gn, hn) and np.array_equal(go, ho)
