"""Pins the oracle's float transcendental functions against libm/numpy.

The product and the oracle use the same published polynomial methods
(Cephes single precision) so their results agree bit for bit; here the
oracle's versions are checked for accuracy against double-precision numpy,
so a wrong coefficient, sign or range reduction fails.
"""
import numpy as np
import pytest

from oracle import orc


def fm(kind, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    orc.lib().orc_fmath(kind, x.ctypes.data, out.ctypes.data, x.size)
    return out


def ulp_error(a, ref):
    ref32 = ref.astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    sp[sp == 0] = np.spacing(np.float32(0))
    return np.abs(a.astype(np.float64) - ref) / sp


@pytest.mark.parametrize("kind,fn,lo,hi,tol", [
    (0, np.sin, -20.0, 20.0, 2.0),
    (1, np.cos, -20.0, 20.0, 2.0),
    (2, np.arccos, -1.0, 1.0, 2.0),
    (3, np.exp, -80.0, 80.0, 2.0),
])
def test_accuracy(kind, fn, lo, hi, tol):
    x = np.linspace(lo, hi, 400001).astype(np.float32)
    err = ulp_error(fm(kind, x), fn(x.astype(np.float64)))
    # sin/cos near their zeros are judged by absolute error
    if kind in (0, 1):
        ref = fn(x.astype(np.float64))
        small = np.abs(ref) < 1e-3
        assert np.abs(fm(kind, x)[small] - ref[small]).max() < 1e-7
        err = err[~small]
    assert err.max() <= tol


def test_log_accuracy():
    x = np.exp(np.linspace(-85, 85, 400001)).astype(np.float32)
    err = ulp_error(fm(4, x), np.log(x.astype(np.float64)))
    near1 = np.abs(x - 1) < 1e-3
    assert err[~near1].max() <= 2.0
    assert np.abs(fm(4, x[near1]) - np.log(x[near1].astype(np.float64))).max() < 2e-8


def test_special_values():
    assert fm(3, [0.0])[0] == 1.0
    assert fm(4, [1.0])[0] == 0.0
    assert fm(3, [-200.0])[0] == 0.0
    assert np.isinf(fm(3, [200.0])[0])
    assert np.isinf(fm(4, [0.0])[0]) and fm(4, [0.0])[0] < 0
    assert np.isnan(fm(4, [-1.0])[0])
    assert fm(2, [1.0])[0] == 0.0
    assert abs(fm(2, [-1.0])[0] - np.float32(np.pi)) <= np.spacing(np.float32(np.pi))
    # denormal argument of log
    d = np.float32(1e-40)
    assert abs(fm(4, [d])[0] - np.log(np.float64(d))) < 1e-5


def test_pythagorean_identity():
    x = np.linspace(-7, 7, 10001).astype(np.float32)
    s, c = fm(0, x).astype(np.float64), fm(1, x).astype(np.float64)
    assert np.abs(s * s + c * c - 1).max() < 4e-7
