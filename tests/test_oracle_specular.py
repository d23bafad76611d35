"""Pins of the OBJ material's specular lobe in the oracle (normalised
Blinn-Phong, bsdfs/Specular.ih:14-78, power-cosine half-vector sampling,
math/sampling.ih:70-87) and of the power function it uses: fpow against
numpy; the sample weight equal to eval value / pdf along the sampled
direction; the sampling pdf's integral over the hemisphere at normal
incidence; and the normalisation's closed form at normal incidence
(the estimate of the reflectance by BSDF sampling and by uniform
hemisphere sampling agree)."""
import ctypes as C

import numpy as np
import pytest

from oracle import orc


def _fpow(x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    f = orc.lib().orc_fpow
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    f(x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out


def _spec(ks, ns, wo, wi, s):
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    s = np.ascontiguousarray(s, np.float32)
    ks = np.ascontiguousarray(ks, np.float32)
    out = np.empty((wo.shape[0], 11), np.float32)
    f = orc.lib().orc_specular
    f.argtypes = [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    f(ks.ctypes.data, ns, wo.ctypes.data, wi.ctypes.data, s.ctypes.data, out.ctypes.data, wo.shape[0])
    return out


def test_fpow_accuracy():
    rng = np.random.default_rng(1)
    x = rng.uniform(1e-3, 1.0, 200000).astype(np.float32)
    y = rng.uniform(0.02, 200.0, 200000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    got = _fpow(x, y).astype(np.float64)
    keep = ref > 1e-30
    rel = np.abs(got[keep] - ref[keep]) / ref[keep]
    # exp(y log x) in float: relative error grows with |y log x| (<= 88 here)
    assert np.all(rel <= 2e-7 * (1.0 + np.abs(y[keep] * np.log(x[keep]))))
    assert _fpow([0.0, 0.0, 1.0, 0.5], [2.0, 0.0, 37.0, 0.0]).tolist() == [0.0, 1.0, 1.0, 1.0]


def _hemi(n, rng):
    u = rng.random((n, 2))
    z = u[:, 0]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2 * np.pi * u[:, 1]
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


@pytest.mark.parametrize("ns", [2.0, 10.0, 60.0])
def test_sample_weight_is_value_over_pdf(ns):
    rng = np.random.default_rng(2)
    n = 20000
    wo = _hemi(n, rng)
    wo[:, 2] = np.maximum(wo[:, 2], 0.05)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    s = rng.random((n, 2))
    ks = (0.6, 0.3, 0.1)
    first = _spec(ks, ns, wo, wo, s)
    wi = first[:, 4:7]
    ok = first[:, 7] > 0
    assert ok.mean() > 0.5
    e = _spec(ks, ns, wo[ok], wi[ok], s[ok])
    value, pdf_e = e[:, 0:3].astype(np.float64), e[:, 3].astype(np.float64)
    pdf_s, weight = first[ok, 7].astype(np.float64), first[ok, 8:11].astype(np.float64)
    good = pdf_e > 1e-3
    assert np.allclose(pdf_e[good], pdf_s[good], rtol=2e-3)
    assert np.allclose(weight[good], value[good] / pdf_e[good, None], rtol=3e-3, atol=1e-6)


@pytest.mark.parametrize("ns", [2.0, 10.0, 60.0])
def test_pdf_integrates_to_one_and_reflectance(ns):
    rng = np.random.default_rng(3)
    n = 400000
    wi = _hemi(n, rng)
    wo = np.tile([0.0, 0.0, 1.0], (n, 1))
    e = _spec((1.0, 1.0, 1.0), ns, wo, wi, np.zeros((n, 2)))
    # uniform hemisphere sampling: pdf 1 / (2 pi).  Half vectors beyond 45
    # degrees reflect normal incidence below the surface: the pdf over the
    # upper hemisphere integrates to 1 - P(cos a < 1/sqrt 2) = 1 - 2^(-(n4+1)/2)
    integral = np.mean(e[:, 3].astype(np.float64)) * 2 * np.pi
    assert abs(integral - (1.0 - 2.0 ** (-(4.0 * ns + 1.0) / 2.0))) < 0.01
    refl_uniform = np.mean(e[:, 0].astype(np.float64)) * 2 * np.pi
    s = rng.random((n, 2))
    smp = _spec((1.0, 1.0, 1.0), ns, wo, wo, s)
    refl_sampled = np.mean(smp[:, 8].astype(np.float64))
    assert abs(refl_uniform - refl_sampled) < 0.02 * max(refl_sampled, 1e-3) + 0.01
    # closed form at normal incidence: wi = reflect(+z, h) with h at angle a
    # has cos(theta_i) = cos(2a); the weight (2 tmp / (n4 + 1)) cos(a)
    # averaged over the power-cosine distribution of cos(a) (n4 = 4 ns)
    n4 = 4.0 * ns
    tmp = (n4 + 2.0) / (2.0 - 2.0 ** (-n4 / 2.0))
    c = np.linspace(0.0, 1.0, 2000001)
    dens = (n4 + 1.0) * c ** n4  # pdf of cos(a) under the power-cosine distribution
    w = np.where(2 * c * c - 1 > 0, 2 * tmp / (n4 + 1.0) * c, 0.0)
    closed = np.trapezoid(dens * w, c)
    assert abs(refl_sampled - closed) < 0.01 * closed + 2e-3
