"""Pins of the oracle's metal and glass BSDFs by closed forms: the
conductor Fresnel term at normal and grazing incidence (bsdfs/Fresnel.ih:
75-84), the GGX distribution's cumulative distribution of half vectors
through the visible-normal sampler and through the eval pdf
(GGXDistribution.ih, MicrofacetConductor.ih:27-110), sample/eval
consistency and the smooth-limit white furnace of the microfacet
conductor, the spectral conductor's resampling of its ior table and its
sum over the spectrum (Metal.cpp:34-57, Fresnel.ih:83-90), and
RobustDielectric's split (RobustDielectric.ih:31-160):
mirror reflection with the probability of the Fresnel equations in angle
form, Snell's law for the refracted direction, total internal reflection
and pass-through at equal indices."""
import ctypes as C

import numpy as np
import pytest

from oracle import orc

SPEC_REFL, GLOSSY_REFL, SPEC_TRANS = 1, 2, 8


def _material(kind, **kv):
    m = orc.OrcMaterial()
    m.type = orc.ORC_MAT[kind]
    m.kd = (C.c_float * 3)(0.8, 0.8, 0.8)
    m.d = 1.0
    m.eta = (C.c_float * 3)(*kv.get("eta", (1.47, 0.984, 0.602)))
    m.k = (C.c_float * 3)(*kv.get("k", (7.64, 6.55, 5.36)))
    m.roughness = kv.get("roughness", 0.1)
    m.etaInside = kv.get("eta_inside", 1.5)
    m.etaOutside = kv.get("eta_outside", 1.0)
    m.attenuationColorInside = (C.c_float * 3)(1.0, 1.0, 1.0)
    m.attenuationColorOutside = (C.c_float * 3)(1.0, 1.0, 1.0)
    m.attenuationDistance = 1.0
    m.color = (C.c_float * 3)(1.0, 1.0, 1.0)
    m.intensity = 1.0
    m.ns = 10.0
    return m


def _metal_ior(table, roughness=0.0):
    m = _material("metal", roughness=roughness)
    t = np.ascontiguousarray(table, np.float32).reshape(-1, 3)
    m.ior = t.ctypes.data
    m.numIor = len(t)
    m._table = t  # keep the array alive with the structure
    return m


# spectrum_sRGB (math/spectrum.ispc:11-38), data of the reference
_SRGB = np.array([
    [0.0598548, -0.0234574, -0.220138, -0.238902, 0.316327, 0.738315, 0.323302, 0.0446981],
    [-0.0567346, -0.0160361, 0.223861, 0.531185, 0.337221, 0.0149718, -0.0296053, -0.00486239],
    [0.420693, 0.616597, 0.0796766, -0.0496266, -0.0473149, -0.0167536, -0.00295686, -0.000314818]])
_WL = 430.0 + 35.0 * np.arange(8)


def test_spectral_conductor_flat_spectrum_is_grey():
    # wavelength-independent eta and k: every sample has the same
    # reflectance F, and F times each row sum of the sRGB weights (all
    # within 2e-6 of 1: a flat spectrum is white) clamps to F or 1
    eta, k = 0.3, 2.9
    m = _metal_ior([[380.0, eta, k], [780.0, eta, k]])
    out = _bsdf(m, np.array([[0.0, 0.0, 1.0]]), (0.0, 0.0, 1.0), (0.5, 0.5), 0.5)
    F = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    np.testing.assert_allclose(out[0, 8:11], np.minimum(1.0, F * _SRGB.sum(1)), rtol=3e-6)
    assert np.all(np.abs(_SRGB.sum(1) - 1.0) < 2e-6)


def test_spectral_conductor_resampling():
    # eta linear in the wavelength between two entries at 400 and 700 nm,
    # k = 0: linear resampling reproduces the line exactly, and the Fresnel
    # term at normal incidence is ((eta - 1) / (eta + 1))^2 per sample
    m = _metal_ior([[400.0, 1.2, 0.0], [700.0, 2.0, 0.0]])
    eta = 1.2 + 0.8 * (_WL - 400.0) / 300.0
    F = ((eta - 1) / (eta + 1)) ** 2
    out = _bsdf(m, np.array([[0.0, 0.0, 1.0]]), (0.0, 0.0, 1.0), (0.5, 0.5), 0.5)
    np.testing.assert_allclose(out[0, 8:11], np.clip(_SRGB @ F, 0.0, 1.0), atol=2e-7)
    # entries only beyond 500 nm (a sloped table, so holding the first entry
    # differs from extrapolating it) and only before 600 nm: the nearest
    # entry holds outside the table; one entry: a flat spectrum
    for table, eta in (([[500.0, 1.2, 0.0], [800.0, 2.0, 0.0]], 1.2 + 0.8 * np.maximum(_WL - 500.0, 0.0) / 300.0),
                       ([[300.0, 1.1, 0.0], [600.0, 1.4, 0.0]], np.minimum(1.1 + 0.3 * (_WL - 300.0) / 300.0, 1.4)),
                       ([[550.0, 1.7, 0.0]], np.full(8, 1.7))):
        F = ((eta - 1) / (eta + 1)) ** 2
        out = _bsdf(_metal_ior(table), np.array([[0.0, 0.0, 1.0]]), (0.0, 0.0, 1.0), (0.5, 0.5), 0.5)
        np.testing.assert_allclose(out[0, 8:11], np.clip(_SRGB @ F, 0.0, 1.0), atol=2e-7)


def _bsdf(m, wo, wi, s, ss):
    """Columns: eval value (3), eval pdf, sample wi (3), sample pdf, sample
    weight (3), sample type, the BSDF's scattering type."""
    n = len(wo)
    wo = np.ascontiguousarray(np.broadcast_to(wo, (n, 3)), np.float32)
    wi = np.ascontiguousarray(np.broadcast_to(wi, (n, 3)), np.float32)
    s = np.ascontiguousarray(np.broadcast_to(s, (n, 2)), np.float32)
    ss = np.ascontiguousarray(np.broadcast_to(ss, (n,)), np.float32)
    out = np.empty((n, 13), np.float32)
    f = orc.lib().orc_bsdf
    f.argtypes = [C.c_void_p] * 6 + [C.c_int64]
    f(C.addressof(m), wo.ctypes.data, wi.ctypes.data, s.ctypes.data, ss.ctypes.data, out.ctypes.data, n)
    return out


def _dir(theta, phi=0.0):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


def _alpha(roughness):
    return max(roughness * roughness, 0.001)  # roughnessToAlpha, MicrofacetDistribution.ih:17-22


def test_conductor_fresnel_closed_forms():
    eta, k = (0.2, 0.45, 1.4), (3.6, 2.4, 1.8)
    m = _material("metal", eta=eta, k=k, roughness=0.0)
    wo = np.array([[0.0, 0.0, 1.0], _dir(np.arccos(1e-3), 0.7)])
    out = _bsdf(m, wo, wo, (0.5, 0.5), 0.5)
    eta, k = np.array(eta), np.array(k)
    # normal incidence: ((eta - 1)^2 + k^2) / ((eta + 1)^2 + k^2)
    np.testing.assert_allclose(out[0, 8:11], ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2), rtol=2e-6)
    # grazing: the reflectance tends to 1; to first order in cos the two
    # polarisations give 1 - 4 eta cos and 1 - 4 eta cos / (eta^2 + k^2)
    # (second-order terms stay below 3e-5 at cos = 1e-3)
    np.testing.assert_allclose(out[1, 8:11], 1.0 - 2.0 * eta * 1e-3 * (1.0 + 1.0 / (eta ** 2 + k ** 2)), atol=3e-5)
    # mirror direction, delta pdf
    np.testing.assert_allclose(out[:, 4:7], wo * np.array([[-1.0, -1.0, 1.0]]), atol=1e-6)
    assert np.all(np.isinf(out[:, 7])) and np.all(out[:, 11] == SPEC_REFL)


def _ggx_cdf(alpha, cos_h):
    """P(theta_h < theta) of the GGX distribution of normals weighted by
    cos theta_h: tan^2 / (alpha^2 + tan^2)."""
    t2 = (1.0 - cos_h ** 2) / cos_h ** 2
    return t2 / (alpha ** 2 + t2)


@pytest.mark.parametrize("roughness", [0.3, 0.5, 0.8])
def test_ggx_sampler_matches_closed_form_cdf(roughness):
    # at normal incidence G1 = 1 and visible-normal sampling draws half
    # vectors with density D cos; wi = reflect(wo, wh) so
    # wi.z = 2 cos^2 theta_h - 1, and no sample leaves the hemisphere iff
    # cos theta_h > sqrt(1/2)
    a = _alpha(roughness)
    n = 200000
    s = np.random.default_rng(3).random((n, 2))
    out = _bsdf(_material("metal", eta=(0.0, 0.0, 0.0), k=(0.0, 0.0, 0.0), roughness=roughness),
                np.array([[0.0, 0.0, 1.0]]).repeat(n, 0), (0.0, 0.0, 1.0), s, 0.5)
    ok = out[:, 7] > 0
    assert abs(ok.mean() - _ggx_cdf(a, np.sqrt(0.5))) < 4.0 * np.sqrt(0.25 / n)
    for t in (0.2, 0.5, 0.8, 0.95):
        frac = np.mean(ok & (out[:, 6] >= t))
        assert abs(frac - _ggx_cdf(a, np.sqrt((1.0 + t) / 2.0))) < 4.0 * np.sqrt(0.25 / n), t
    # F = 1 at eta = k = 0: the sample weight is G2 / G1 <= 1
    assert np.all(out[ok, 8:11] <= 1.0 + 1e-6) and np.all(out[ok, 8:11] > 0.0)


@pytest.mark.parametrize("roughness", [0.3, 0.5, 0.8])
def test_ggx_eval_pdf_matches_closed_form_cdf(roughness):
    # integral of the eval pdf over the cap wi.z >= t (normal incidence,
    # rotationally symmetric): 2 pi int_t^1 pdf dz
    a = _alpha(roughness)
    m = _material("metal", eta=(0.0, 0.0, 0.0), k=(0.0, 0.0, 0.0), roughness=roughness)
    for t in (0.0, 0.5, 0.9):
        x = np.linspace(0.0, 1.0, 40001)
        z = 1.0 - (1.0 - t) * x * x  # dense near the peak at z = 1
        wi = np.stack([np.sqrt(np.maximum(0.0, 1.0 - z * z)), np.zeros_like(z), z], 1)
        pdf = _bsdf(m, np.array([[0.0, 0.0, 1.0]]).repeat(len(z), 0), wi, (0.5, 0.5), 0.5)[:, 3].astype(np.float64)
        integral = 2.0 * np.pi * np.trapezoid(pdf * 2.0 * (1.0 - t) * x, x)
        assert integral == pytest.approx(_ggx_cdf(a, np.sqrt((1.0 + t) / 2.0)), rel=2e-3), t


@pytest.mark.parametrize("theta", [0.4, 1.1])
def test_microfacet_conductor_sample_eval_consistency(theta):
    rng = np.random.default_rng(5)
    n = 20000
    wo = _dir(theta, 0.3)
    m = _material("metal", eta=(0.2, 0.45, 1.4), k=(3.6, 2.4, 1.8), roughness=0.45)
    out = _bsdf(m, wo[None].repeat(n, 0), (0.0, 0.0, 1.0), rng.random((n, 2)), 0.5)
    ok = out[:, 7] > 0
    assert ok.mean() > 0.5 and np.all(out[ok, 11] == GLOSSY_REFL)
    ev = _bsdf(m, wo[None].repeat(ok.sum(), 0), out[ok, 4:7], (0.5, 0.5), 0.5)
    np.testing.assert_allclose(ev[:, 3], out[ok, 7], rtol=2e-4)
    np.testing.assert_allclose(ev[:, 0:3] / ev[:, 3:4], out[ok, 8:11], rtol=5e-4)
    # the fraction of samples that stay above the horizon equals the integral
    # of the eval pdf over the hemisphere
    zs = (np.arange(1200) + 0.5) / 1200
    ph = (np.arange(1200) + 0.5) / 1200 * 2 * np.pi
    Z, P = np.meshgrid(zs, ph, indexing="ij")
    r = np.sqrt(1 - Z * Z)
    wi = np.stack([r * np.cos(P), r * np.sin(P), Z], -1).reshape(-1, 3)
    pdf = _bsdf(m, wo[None].repeat(len(wi), 0), wi, (0.5, 0.5), 0.5)[:, 3].astype(np.float64)
    integral = pdf.sum() * (1.0 / 1200) * (2 * np.pi / 1200)
    assert integral == pytest.approx(ok.mean(), abs=0.01)


@pytest.mark.parametrize("theta", [0.5, 1.0])
def test_microfacet_conductor_smooth_limit_white_furnace(theta):
    # F = 1 (eta = k = 0) and alpha at its floor 0.001: shadowing vanishes,
    # so every sample reflects all energy
    n = 20000
    m = _material("metal", eta=(0.0, 0.0, 0.0), k=(0.0, 0.0, 0.0), roughness=0.01)
    out = _bsdf(m, _dir(theta)[None].repeat(n, 0), (0.0, 0.0, 1.0), np.random.default_rng(7).random((n, 2)), 0.5)
    assert np.mean(out[:, 8]) == pytest.approx(1.0, abs=2e-3)


def _fresnel_angles(cos_i, n):
    """Fresnel reflectance of unpolarised light from the angle form
    Rs = sin^2(ti - tt) / sin^2(ti + tt), Rp = tan^2(ti - tt) / tan^2(ti + tt)."""
    ti = np.arccos(cos_i)
    tt = np.arcsin(np.sin(ti) / n)
    rs = np.sin(ti - tt) ** 2 / np.sin(ti + tt) ** 2
    rp = np.tan(ti - tt) ** 2 / np.tan(ti + tt) ** 2
    return 0.5 * (rs + rp)


@pytest.mark.parametrize("theta", [0.01, 0.3, 0.8, 1.2, 1.5])
def test_dielectric_fresnel_split_and_snell(theta):
    n = 4096
    ss = (np.arange(n) + 0.5) / n
    wo = _dir(theta, 1.0)
    out = _bsdf(_material("glass"), wo[None].repeat(n, 0), (0.0, 0.0, 1.0), (0.5, 0.5), ss)
    refl = out[:, 11] == SPEC_REFL
    assert np.all(refl | (out[:, 11] == SPEC_TRANS))
    assert refl.mean() == pytest.approx(_fresnel_angles(np.cos(theta), 1.5), abs=2.0 / n)
    assert np.all(np.isinf(out[:, 7])) and np.all(out[:, 8:11] == 1.0)
    np.testing.assert_allclose(out[refl, 4:7], np.broadcast_to(wo * np.array([-1.0, -1.0, 1.0]), (refl.sum(), 3)),
                               atol=4e-6)
    t = out[~refl, 4:7]
    sin_t = np.hypot(t[:, 0], t[:, 1])
    np.testing.assert_allclose(sin_t, np.sin(theta) / 1.5, atol=2e-6)  # Snell's law
    np.testing.assert_allclose(t[:, 2], -np.sqrt(1 - (np.sin(theta) / 1.5) ** 2), atol=2e-6)
    # in the plane of incidence, on the far side of the normal
    np.testing.assert_allclose(t[:, :2] / sin_t[:, None], np.broadcast_to(-wo[:2] / np.sin(theta), (len(t), 2)),
                               atol=2e-5)


def test_dielectric_total_internal_reflection_and_index_match():
    n = 512
    ss = (np.arange(n) + 0.5) / n
    # from the denser side (outside 1.5, inside 1.0): beyond asin(1/1.5) every sample reflects
    for theta, expect in ((0.9, 1.0), (0.3, _fresnel_angles(np.cos(0.3), 1.0 / 1.5))):
        out = _bsdf(_material("glass", eta_inside=1.0, eta_outside=1.5), _dir(theta)[None].repeat(n, 0),
                    (0.0, 0.0, 1.0), (0.5, 0.5), ss)
        assert (out[:, 11] == SPEC_REFL).mean() == pytest.approx(expect, abs=2.0 / n)
    # equal indices: no reflection, the ray passes straight through
    wo = _dir(0.7, 2.0)
    out = _bsdf(_material("glass", eta_inside=1.3, eta_outside=1.3), wo[None].repeat(n, 0), (0.0, 0.0, 1.0),
                (0.5, 0.5), ss)
    assert np.all(out[:, 11] == SPEC_TRANS)
    np.testing.assert_allclose(out[:, 4:7], -wo[None].repeat(n, 0), atol=2e-6)
