"""BSDFs, Fresnel, Burley profile and light sampling of the oracle, pinned by
closed forms (reference: Src/Core/material.cu)."""
import math

import numpy as np
import pytest
from scipy import integrate, stats


def _mat(type_, base=(0.8, 0.6, 0.4), rough=0.5, metal=0.0, ior=1.5, trans=1.0, scatter=(0.2, 0.1, 0.05)):
    m = np.zeros(16, np.float32)
    m[0] = type_
    m[1:4] = base
    m[7] = rough
    m[8] = metal
    m[9] = trans
    m[10] = ior
    m[11:14] = scatter
    return m


def _v(*x):
    return np.array(x, np.float32)


def _dir(theta, phi):
    return _v(math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta))


def test_fresnel_closed_forms(oracle_lib):
    L = oracle_lib.lib()
    assert L.ref_fresnel_dielectric(1.0, 1.0, 1.5) == pytest.approx(0.04, abs=1e-7)
    assert L.ref_fresnel_dielectric(-1.0, 1.0, 1.5) == pytest.approx(0.04, abs=1e-7)  # leaving: eta swapped
    # Total internal reflection from inside beyond the critical angle asin(1/1.5).
    assert L.ref_fresnel_dielectric(-math.cos(math.radians(50)), 1.0, 1.5) == 1.0
    assert L.ref_fresnel_dielectric(-math.cos(math.radians(40)), 1.0, 1.5) < 1.0
    assert L.ref_fresnel_dielectric(1e-6, 1.0, 1.5) == pytest.approx(1.0, abs=1e-3)  # grazing


@pytest.mark.parametrize("rough", [0.2, 0.45, 0.8, 1.0])
def test_ggx_normalisation(oracle_lib, rough):
    """int D(h) (n.h) dw_h = 1 for the GGX distribution with alpha = roughness^2
    (up to the reference's PI = 3.1415926)."""
    L = oracle_lib.lib()
    m = _mat(2, rough=rough)
    n = _v(0, 0, 1)

    def f(theta):
        h = _dir(theta, 0.3)
        d = L.ref_ggx_d(m.ctypes.data, h.ctypes.data, n.ctypes.data, h.ctypes.data)
        return d * math.cos(theta) * math.sin(theta) * 2 * math.pi

    a2 = (rough * rough) ** 2
    pts = [math.atan(math.sqrt(a2) * k) for k in (0.5, 1, 2, 4)]
    val, _ = integrate.quad(f, 0, math.pi / 2, points=pts, limit=400)
    assert val == pytest.approx(math.pi / 3.1415926, rel=2e-4)


def test_diffuse_sampling_is_cosine_weighted(oracle_lib):
    """cos^2(theta) of theta = acos(1 - 2 u) / 2 is uniform; the reported pdf is cos/PI."""
    L = oracle_lib.lib()
    m = _mat(1)
    n = _v(0.0, 0.6, 0.8)
    wi = _v(0, 0, 1)
    rng = np.random.default_rng(1)
    c2 = []
    out = np.zeros(3, np.float32)
    for u in rng.random((20000, 3)).astype(np.float32):
        pdf = L.ref_bsdf_sample(m.ctypes.data, wi.ctypes.data, n.ctypes.data, u.ctypes.data, out.ctypes.data)
        c = float(np.dot(out, n))
        assert pdf == pytest.approx(max(c, 0) / 3.1415926, abs=1e-6)
        c2.append(c * c)
    assert stats.kstest(c2, "uniform").pvalue > 1e-3


@pytest.mark.parametrize("rough,metal", [(0.3, 0.0), (0.6, 1.0), (0.9, 0.0)])
def test_pbr_pdf_matches_its_sampler(oracle_lib, rough, metal):
    """E[1{n.wo > 0} / pdf(wo)] over the sampler's own draws is the solid angle
    of the upper hemisphere (2 pi) when pdf is the sampler's density there."""
    L = oracle_lib.lib()
    m = _mat(2, rough=rough, metal=metal)
    n = _v(0, 0, 1)
    wi = _dir(0.6, 1.0)
    rng = np.random.default_rng(int(rough * 10))
    out = np.zeros(3, np.float32)
    acc = []
    for u in rng.random((40000, 3)).astype(np.float32):
        pdf = L.ref_bsdf_sample(m.ctypes.data, wi.ctypes.data, n.ctypes.data, u.ctypes.data, out.ctypes.data)
        acc.append(1.0 / pdf if out[2] > 0 and pdf > 0 else 0.0)
    est = np.mean(acc)
    err = np.std(acc) / math.sqrt(len(acc))
    assert abs(est - 2 * math.pi) < 4 * err + 0.01 * 2 * math.pi


def test_pbr_eval_reciprocity_and_hemisphere(oracle_lib):
    L = oracle_lib.lib()
    m = _mat(2, rough=0.5, metal=0.3)
    n = _v(0, 0, 1)
    a, b = _dir(0.4, 0.2), _dir(1.1, 2.0)
    f1, f2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    L.ref_bsdf_eval(m.ctypes.data, a.ctypes.data, n.ctypes.data, b.ctypes.data, f1.ctypes.data)
    L.ref_bsdf_eval(m.ctypes.data, b.ctypes.data, n.ctypes.data, a.ctypes.data, f2.ctypes.data)
    assert np.allclose(f1, f2, rtol=1e-5)
    below = _v(0.3, 0, -0.95)
    pdf = L.ref_bsdf_eval(m.ctypes.data, a.ctypes.data, n.ctypes.data, below.ctypes.data, f1.ctypes.data)
    assert (f1 == 0).all() and pdf == 0.0


def test_dielectric_normal_incidence(oracle_lib):
    L = oracle_lib.lib()
    m = _mat(3, base=(0.5, 0.7, 0.9), trans=0.5, ior=1.5)
    n = _v(0, 0, 1)
    wi = _v(0, 0, 1)  # towards the viewer, on the normal side
    out = np.zeros(3, np.float32)
    pdf = L.ref_bsdf_sample(m.ctypes.data, wi.ctypes.data, n.ctypes.data, _v(0.01, 0, 0).ctypes.data, out.ctypes.data)
    assert np.allclose(out, [0, 0, 1]) and pdf == pytest.approx(0.04, abs=1e-6)   # reflected, pdf = F
    pdf = L.ref_bsdf_sample(m.ctypes.data, wi.ctypes.data, n.ctypes.data, _v(0.5, 0, 0).ctypes.data, out.ctypes.data)
    assert np.allclose(out, [0, 0, -1]) and pdf == pytest.approx(0.96, abs=1e-6)  # transmitted, pdf = 1 - F


def test_burley_profile_normalisation(oracle_lib):
    """int_0^inf R(r) 2 pi r dr = 1 for (e^{-r/d} + e^{-r/3d}) / (8 pi d r)."""
    L = oracle_lib.lib()
    m = _mat(4, scatter=(0.3, 0.1, 0.02))
    for ch, d in enumerate((0.3, 0.1, 0.02)):
        val, _ = integrate.quad(lambda r: L.ref_burley(m.ctypes.data, r, ch) * 2 * math.pi * r, 0, 60 * d, limit=400)
        assert val == pytest.approx(math.pi / 3.1415926, rel=1e-4)
    assert L.ref_burley(m.ctypes.data, 0.0, 0) == 0.0
