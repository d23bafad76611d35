"""Pins of the neural-radiance-cache oracle (oracle/nrc_ref.py): the
backward pass against central finite differences in FP64, the bfloat16
rounding against its definition, and the frequency encoding against
closed forms."""
import numpy as np

from oracle import nrc_ref as R


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = R.init_params(5)
    x = rng.uniform(-1, 1, (6, R.N_IN))
    t = rng.uniform(0, 2, (6, 3))
    _, g = R.loss_and_grad(params, x, t)
    y0, _ = R.forward(params, x)
    den = y0 * y0 + R.LOSS_EPS  # the gradient holds the denominator constant

    def L(p):
        y, _ = R.forward(p, x)
        return np.mean((y - t) ** 2 / den)

    idx = np.concatenate([rng.choice(R.OFFSETS[i + 1] - R.OFFSETS[i], 12, replace=False) + R.OFFSETS[i]
                          for i in range(5)])
    idx = idx[(idx < R.OFFSETS[4] + 3 * 64) | (idx >= R.OFFSETS[4] + 16 * 64)]  # skip padded output rows
    h = 1e-6
    for i in idx:
        e = np.zeros_like(params)
        e[i] = h
        fd = (L(params + e) - L(params - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])
    # padded output rows get no gradient
    assert np.all(g[R.OFFSETS[4] + 3 * 64:] == 0)


def test_bf16_round_to_nearest_even():
    # bfloat16 keeps 8 significant bits: spacing 2^-7 just above 1
    vals = np.array([1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -12, -2.5, 65504.0],
                    np.float32)
    got = R.bf16_round(vals)
    assert got[0] == 1.0 + 2 ** -7  # representable
    assert got[1] == 1.0  # tie -> even
    assert got[2] == 1.0 + 2 ** -6  # tie -> even
    assert got[3] == 1.0 + 2 ** -7  # above the tie
    assert got[4] == -2.5
    assert np.all(np.abs(got - vals) <= np.abs(vals) * 2 ** -8)


def test_encoding_closed_forms():
    raw = np.array([[0.5, -1.0, 3.0, 0.0, 0.6, 0.8, 0.0, 1.0, 0.0, 0.2, 0.4, 0.6]])
    x = R.encode(raw, (-1.0, -1.0, -1.0), (1.0, 1.0, 5.0))
    p = np.array([0.75, 0.0, 4.0 / 6.0])
    for c, v in enumerate(list(p) + [0.0, 0.6, 0.8]):
        for k in range(R.N_BANDS):
            a = (2 ** k) * np.pi * v
            assert abs(x[0, c * 24 + 2 * k] - np.sin(a)) < 1e-12
            assert abs(x[0, c * 24 + 2 * k + 1] - np.cos(a)) < 1e-12
    np.testing.assert_allclose(x[0, 144:147], [0.2, 0.4, 0.6])
    np.testing.assert_allclose(x[0, 147:150], [0.0, 1.0, 0.0])
    np.testing.assert_allclose(x[0, 150:153], p)
    np.testing.assert_allclose(x[0, 153:156], [0.0, 0.6, 0.8])
    assert x[0, 156] == 1.0 and np.all(x[0, 157:] == 0)
    assert R.N_PARAMS == 23552 and R.OFFSETS[:5] == [0, 10240, 14336, 18432, 22528]
