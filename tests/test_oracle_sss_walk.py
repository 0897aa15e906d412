"""The oracle's random-walk subsurface transport, pinned by what the physics
fixes (the walk is this library's addition; the reference's subsurface
material is the Burley profile, so there is no reference vector for it).

* Beer-Lambert: a purely absorbing slab (sigma_s = 0) entered in a
  cosine-distributed direction transmits 2 E3(sigma_a d) per channel (E3 the
  exponential integral of order 3).  Catches a wrong free-flight pdf or exit
  weight (T instead of T / mean T, a channel's T for all), a dropped or
  reversed entry direction, a wrong sign in the transmittance.
* Chandrasekhar: a half-space of isotropic scatterers, index-matched, lit
  diffusely, reflects 1 - 2 sqrt(1 - albedo) * integral of H(mu) mu over
  [0, 1], with H the solution of H(mu) = 1 / (1 - albedo/2 * mu * integral of
  H(mu') / (mu + mu') dmu') -- a property of the albedo alone, so three
  channels of different sigma_t and albedo each meet their own value in one
  run.  Catches an anisotropic or mis-normalised phase sample, a scattering
  weight of sigma_s / sigma_t of one channel instead of the mixture, a biased
  roulette.
* White furnace: with sigma_a = 0 and grey sigma_s every walk that leaves has
  weight 1 and almost every walk leaves a small sphere.
* Exit geometry: the exit point lies on the surface, its normal faces out.
"""
import numpy as np
import pytest
from scipy import special

from cudaraytracer_amd import scenes
from oracle import rtref


def _walk_scene(tmp_path, mesh, sigma_a, sigma_s):
    spec = scenes.SceneSpec(
        camera=scenes._camera(8, 8, 1, 2),
        materials=[{"name": "medium", "type": "subsurface", "basecolor": [1, 1, 1], "ior": 1.0,
                    "subsurface_method": "random_walk", "sigma_a": list(map(float, sigma_a)),
                    "sigma_s": list(map(float, sigma_s))}],
        meshes=[scenes.MeshSpec("body", "medium", mesh)],
        lights=[])
    a = rtref.parse_scene(scenes.write_scene(spec, str(tmp_path), "walk"))
    return rtref.RefScene(a)


def _uniforms(n, seed):
    return np.random.default_rng(seed).random((n, 2), np.float32) * np.float32(0.999999) + np.float32(1e-7)


def _chandrasekhar_albedo(albedo: float, n: int = 400) -> float:
    """Spherical albedo of an isotropically scattering half-space (index-matched
    boundary, diffuse illumination) from Chandrasekhar's H-function, solved by
    fixed-point iteration of its equivalent form 1 / H(mu) = sqrt(1 - albedo) +
    albedo/2 * integral of mu' H(mu') / (mu + mu') dmu' on a Gauss-Legendre grid
    over [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    mu, wt = 0.5 * (x + 1), 0.5 * w
    H = np.ones(n)
    for _ in range(5000):
        Hn = 1.0 / (np.sqrt(1 - albedo) + 0.5 * albedo * ((wt * mu * H)[None, :] / (mu[:, None] + mu[None, :])).sum(1))
        if np.max(np.abs(Hn - H)) < 1e-14:
            H = Hn
            break
        H = Hn
    # zeroth moment has a closed form: (2 / albedo) (1 - sqrt(1 - albedo))
    assert (wt * H).sum() == pytest.approx(2 / albedo * (1 - np.sqrt(1 - albedo)), rel=1e-6)
    return 1.0 - 2.0 * np.sqrt(1.0 - albedo) * (wt * H * mu).sum()


def test_chandrasekhar_reference_values():
    # The albedo grows with the single-scattering albedo, and near conservative
    # scattering follows 1 - 4/sqrt(3) sqrt(1 - albedo) (H's first moment tends
    # to 2/sqrt(3)).
    vals = [_chandrasekhar_albedo(a) for a in (0.3, 0.6, 0.9, 0.99, 0.999)]
    assert all(a < b for a, b in zip(vals, vals[1:]))
    assert vals[-1] == pytest.approx(1 - 4 / np.sqrt(3) * np.sqrt(1e-3), abs=5e-3)
    assert 0 < vals[0] < 0.3 / 2


def test_walk_beer_lambert_slab(tmp_path, oracle_lib):
    d, half = 1.0, 50.0
    sigma_a = np.array([0.5, 1.0, 2.0])
    rs = _walk_scene(tmp_path, scenes._box((-half, -d, -half), (half, 0.0, half)), sigma_a, [0, 0, 0])
    n = 400_000
    ok, op, on, w, steps = rs.sss_walks(0, (0, 0, 0), (0, 1, 0), _uniforms(n, 1))
    expect = 2 * special.expn(3, sigma_a * d)
    mean = w.astype(np.float64).mean(0)
    se = w.astype(np.float64).std(0) / np.sqrt(n)
    assert np.all(np.abs(mean - expect) < 5 * se + 1e-4), (mean, expect, se)
    # every surviving walk crossed the slab in one segment and left through the bottom
    assert np.all(steps[ok] == 1)
    np.testing.assert_allclose(op[ok][:, 1], -d, atol=1e-5)
    np.testing.assert_array_equal(on[ok], np.tile(np.float32([0, -1, 0]), (ok.sum(), 1)))
    # weight of one exit = T / mean(T) for its own chord (the walk starts 1e-4 inside)
    chord = np.linalg.norm(op[ok].astype(np.float64) - [0, -1e-4, 0], axis=1)[:, None]
    T = np.exp(-sigma_a[None, :] * chord)
    np.testing.assert_allclose(w[ok], T / T.mean(1, keepdims=True), rtol=5e-5)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_walk_half_space_albedo_matches_chandrasekhar(tmp_path, oracle_lib, axis):
    # The half-space's surface normal along each world axis in turn, so a phase
    # sample that favours one axis shows in one of the runs.
    sigma_t = np.array([0.5, 1.0, 2.0])
    albedo = np.array([0.9, 0.6, 0.3])
    sigma_s = albedo * sigma_t
    sigma_a = sigma_t - sigma_s
    half = 500.0
    lo, hi = np.full(3, -half), np.full(3, half)
    lo[axis], hi[axis] = -200.0, 0.0
    rs = _walk_scene(tmp_path, scenes._box(tuple(lo), tuple(hi)), sigma_a, sigma_s)
    up = np.zeros(3, np.float32)
    up[axis] = 1
    n = 300_000
    ok, op, on, w, steps = rs.sss_walks(0, (0, 0, 0), up, _uniforms(n, 2 + axis))
    top = ok & (op[:, axis] > -1e-3)
    assert np.array_equal(top, ok)  # nothing reaches the far faces
    expect = np.array([_chandrasekhar_albedo(a) for a in albedo])
    ww = w.astype(np.float64)
    mean, se = ww.mean(0), ww.std(0) / np.sqrt(n)
    assert np.all(np.abs(mean - expect) < 5 * se), (mean, expect, se)
    assert np.all(on[ok] == up)
    assert steps.max() <= 256


def test_walk_white_furnace_sphere(tmp_path, oracle_lib):
    V, F = scenes.icosphere(3)
    R = 1.0
    rs = _walk_scene(tmp_path, (V[F] * R).astype(np.float32), [0, 0, 0], [2.0, 2.0, 2.0])
    tri = (V[F][0] * R)
    p = tri.mean(0).astype(np.float32)
    nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
    n = 50_000
    ok, op, on, w, steps = rs.sss_walks(0, p, nrm, _uniforms(n, 3))
    assert ok.mean() > 0.999
    np.testing.assert_allclose(w[ok], 1.0, rtol=1e-5)
    r = np.linalg.norm(op[ok], axis=1)
    assert np.all(np.abs(r - R) < 5e-3)  # on the tessellated sphere
    assert np.all(np.einsum("ij,ij->i", on[ok], op[ok] / r[:, None]) > 0.98)  # facing out
    assert 1 < steps[ok].mean() < 64


def test_walk_chromatic_furnace_is_unbiased(tmp_path, oracle_lib):
    # sigma_a = 0: each channel's expected weight is 1 (every photon leaves); the
    # walks a roulette ends are part of the estimator, not a loss.  A moderate
    # spread of sigma_s keeps the one-sample spectral weights light-tailed
    # enough for a standard-error bound (at a 12x spread the bluest channel's
    # weights reach 1e5 and its sample mean converges slowly, unbiased all the same).
    V, F = scenes.icosphere(3)
    rs = _walk_scene(tmp_path, V[F].astype(np.float32), [0, 0, 0], [1.0, 1.5, 2.5])
    tri = V[F][5]
    p = tri.mean(0).astype(np.float32)
    nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
    n = 400_000
    ok, op, on, w, steps = rs.sss_walks(0, p, nrm, _uniforms(n, 4))
    ww = w.astype(np.float64)
    mean, se = ww.mean(0), ww.std(0) / np.sqrt(n)
    assert np.all(np.abs(mean - 1.0) < 5 * se), (mean, se)
    assert steps.max() < 256


def test_walk_is_deterministic_per_coordinate(tmp_path, oracle_lib):
    V, F = scenes.icosphere(2)
    rs = _walk_scene(tmp_path, V[F].astype(np.float32), [0.2, 0.3, 0.4], [1.0, 2.0, 3.0])
    tri = V[F][0]
    p = tri.mean(0).astype(np.float32)
    nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0]).astype(np.float32)
    nrm /= np.linalg.norm(nrm)
    a = rs.sss_walk(0, p, nrm, 0.3, 0.7, seed=9, pixel=11, sample=2, bounce=3)
    b = rs.sss_walk(0, p, nrm, 0.3, 0.7, seed=9, pixel=11, sample=2, bounce=3)
    c = rs.sss_walk(0, p, nrm, 0.3, 0.7, seed=9, pixel=11, sample=2, bounce=4)
    assert a[0] == b[0] and a[4] == b[4]
    for x, y in zip(a[1:4], b[1:4]):
        assert np.array_equal(x, y)
    assert not (np.array_equal(a[1], c[1]) and a[4] == c[4])


@pytest.mark.parametrize("bad,match", [
    ({"subsurface_method": "diffusion"}, "subsurface_method"),
    ({"subsurface_method": "random_walk", "sigma_a": [-1, 0, 0]}, "sigma"),
    ({"subsurface_method": "random_walk", "sigma_a": [0, 0, 0], "sigma_s": [1, 0, 1]}, "sigma"),
    ({"subsurface_method": "burley", "sigma_s": [1e39, 1, 1]}, "sigma"),  # overflows float
])
def test_bad_medium_keys_fail_the_load(tmp_path, bad, match):
    import cudaraytracer_amd as rt

    mat = {"name": "medium", "type": "subsurface", "basecolor": [1, 1, 1]}
    mat.update(bad)
    spec = scenes.SceneSpec(camera=scenes._camera(8, 8, 1, 2), materials=[mat],
                            meshes=[scenes.MeshSpec("body", "medium", scenes._box((0, 0, 0), (1, 1, 1)))],
                            lights=[])
    path = scenes.write_scene(spec, str(tmp_path), "bad")
    with pytest.raises(rt.RtError, match=match):
        rt.Scene(path, bvh_builder=rt.BVH_HOST)


def test_medium_keys_load_with_defaults(tmp_path):
    import cudaraytracer_amd as rt

    mats = [{"name": "rw", "type": "subsurface", "basecolor": [1, 1, 1], "subsurface_method": "random_walk"},
            {"name": "plain", "type": "subsurface", "basecolor": [1, 1, 1]}]
    spec = scenes.SceneSpec(camera=scenes._camera(8, 8, 1, 2), materials=mats,
                            meshes=[scenes.MeshSpec("a", "rw", scenes._box((0, 0, 0), (1, 1, 1))),
                                    scenes.MeshSpec("b", "plain", scenes._box((2, 0, 0), (3, 1, 1)))],
                            lights=[])
    path = scenes.write_scene(spec, str(tmp_path), "ok")
    m = rt.Scene(path, bvh_builder=rt.BVH_HOST).materials()
    ints = m.view(np.int32)
    assert ints[0, 22] == 1 and ints[1, 22] == 0  # random walk, Burley
    assert np.array_equal(m[0, 16:22], np.float32([0, 0, 0, 1, 1, 1]))  # sigma_a, sigma_s defaults
    a = rtref.parse_scene(path)  # the oracle's own parser reads the same keys
    assert np.array_equal(a.materials[:, 16:22], m[:, 16:22])
    assert list(a.materials[:, 22]) == [1.0, 0.0]
