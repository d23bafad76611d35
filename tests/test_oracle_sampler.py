"""Pins the oracle's integer sampling machinery (random.ih, sobol.ih).

Checks are against published known answers and against structural
properties a plausible mistake would break -- not against a re-typing of the
same formulas:
  * PCG32 XSH-RR: O'Neill's pcg32-demo output for seed 42, sequence 54;
  * MurmurHash3_mix/finalize: MurmurHash3_x86_32 test vectors of the
    SMHasher suite, composed from the oracle's mix and finalize steps;
  * Sobol generator table: entries of modules/cpu/math/sobol.inl:33 and the
    (0,1)-sequence property of every dimension plus the (0,m,2)-net property
    of dimensions (0,1);
  * Owen scrambling (OwenHash2 in LDSampler): the scrambled pixel samples of
    one pixel stay a (0,m,2)-net (nested uniform scrambling preserves
    elementary-interval stratification).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import orc


def _hash(kind, xs, seeds=None):
    x = np.ascontiguousarray(xs, np.uint32)
    s = None if seeds is None else np.ascontiguousarray(seeds, np.uint32)
    out = np.empty_like(x)
    orc.lib().orc_hash(kind, x.ctypes.data, None if s is None else s.ctypes.data, out.ctypes.data, x.size)
    return out


def test_pcg32_known_answers():
    # pcg32-demo: pcg32_srandom_r(&rng, 42u, 54u) -> 0xa15c02b7 0x7b47f409 ...
    mult, inc = 6364136223846793005, (54 << 1) | 1
    state = ((0 * mult + inc) + 42) % 2**64
    state = (state * mult + inc) % 2**64
    out = (C.c_uint32 * 6)()
    orc.lib().orc_pcg_raw(C.c_uint64(state), C.c_uint32(inc), out, 6)
    assert list(out) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def _murmur3_x86_32(data: bytes, seed: int) -> int:
    h = seed
    n = len(data)
    for i in range(0, n - n % 4, 4):
        k = int.from_bytes(data[i:i + 4], "little")
        h = int(_hash(0, [h], [k])[0])
    tail = data[n - n % 4:]
    if tail:
        k1 = int.from_bytes(tail, "little")
        k1 = (k1 * 0xCC9E2D51) & 0xFFFFFFFF
        k1 = ((k1 << 15) | (k1 >> 17)) & 0xFFFFFFFF
        k1 = (k1 * 0x1B873593) & 0xFFFFFFFF
        h ^= k1
    h ^= n
    return int(_hash(1, [h])[0])


@pytest.mark.parametrize("data,seed,expect", [
    (b"", 0, 0x00000000),
    (b"", 1, 0x514E28B7),
    (b"", 0xFFFFFFFF, 0x81F16F39),
    (b"\0\0\0\0", 0, 0x2362F9DE),
    (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"Hello, world!", 0x9747B28C, 0x24884CBA),
    (b"The quick brown fox jumps over the lazy dog", 0x9747B28C, 0x2FA826CD),
])
def test_murmur3_known_answers(data, seed, expect):
    assert _murmur3_x86_32(data, seed) == expect


def _sobol_table():
    t = (C.c_uint32 * 128)()
    orc.lib().orc_sobol_rev_table(t)
    return np.array(list(t), np.uint64).reshape(32, 4)


def test_sobol_table_spot_values():
    # modules/cpu/math/sobol.inl:33-161 (bit-reversed, interleaved rows)
    T = _sobol_table()
    assert list(T[0]) == [0x1, 0x1, 0x1, 0x1]
    assert list(T[1]) == [0x3, 0x3, 0x3, 0x2]
    assert list(T[3]) == [0xF, 0x9, 0xA, 0xD]
    assert list(T[9]) == [0x303, 0x339, 0x2A4, 0x3F4]
    assert list(T[22]) == [0x550055, 0x717777, 0x79E7DB, 0x6DF05A]
    assert list(T[31]) == [0xFFFFFFFF, 0xAAAA00A3, 0xC900A00A, 0xB9DBB00D]


def _sobol_points(m):
    """Unscrambled points of dims 0..4 for indices 0..2^m-1 as integers in [0, 2^32)."""
    T = _sobol_table()
    n = 1 << m
    idx = np.arange(n, dtype=np.uint64)
    rev = _hash(5, idx.astype(np.uint32)).astype(np.uint64)
    pts = np.zeros((n, 5), np.uint64)
    pts[:, 0] = rev  # dimension 0 (van der Corput) is the reversed index
    acc = np.zeros((n, 4), np.uint64)
    r = rev.copy()
    for k in range(32):
        bit = (r >> np.uint64(31 - k)) & np.uint64(1)
        acc ^= bit[:, None] * T[k][None, :]
    for d in range(4):
        pts[:, d + 1] = _hash(5, acc[:, d].astype(np.uint32)).astype(np.uint64)
    return pts


@pytest.mark.parametrize("m", [4, 8, 10])
def test_sobol_dimensions_are_stratified(m):
    pts = _sobol_points(m)
    for d in range(5):
        cells = pts[:, d] >> np.uint64(32 - m)
        assert np.array_equal(np.sort(cells), np.arange(1 << m)), f"dim {d} not a (0,1)-net"


@pytest.mark.parametrize("m", [4, 6, 8])
def test_sobol_dims01_are_a_0m2_net(m):
    pts = _sobol_points(m)
    for a in range(m + 1):
        b = m - a
        cx = pts[:, 0] >> np.uint64(32 - a) if a else np.zeros(len(pts), np.uint64)
        cy = pts[:, 1] >> np.uint64(32 - b) if b else np.zeros(len(pts), np.uint64)
        cells = cx * np.uint64(1 << b) + cy
        assert len(np.unique(cells)) == 1 << m, f"elementary intervals {a}x{b} not stratified"


@pytest.mark.parametrize("seed", [0x49C5F37E, 0x12345678, 0xDEADBEEF])
def test_scrambled_pixel_samples_stay_a_net(seed):
    m = 6
    n = 1 << m
    out = np.empty((n, 5), np.float32)
    L = orc.lib()
    for i in range(n):
        buf = (C.c_float * 5)()
        L.orc_ld_samples5(seed, i, 16, buf)
        out[i] = list(buf)
    # 1D stratification of every dimension and 2D of the pixel pair
    for d in range(2):
        cells = np.floor(out[:, d] * n).astype(int)
        assert np.array_equal(np.sort(cells), np.arange(n))
    for a in range(m + 1):
        b = m - a
        cells = np.floor(out[:, 0] * (1 << a)).astype(int) * (1 << b) + np.floor(out[:, 1] * (1 << b)).astype(int)
        assert len(np.unique(cells)) == n


def test_owen_hash_is_nested():
    # in the reversed domain, output bit k depends only on input bits <= k
    rng = np.random.default_rng(0)
    x = rng.integers(0, 2**32, 4096, dtype=np.uint64).astype(np.uint32)
    seeds = np.full(x.size, 0x9374B0D3, np.uint32)
    hx = _hash(3, x, seeds)
    for k in (1, 5, 12, 20, 31):
        mask = np.uint32((1 << k) - 1)
        y = (x & mask) | (rng.integers(0, 2**32, x.size, dtype=np.uint64).astype(np.uint32) & ~mask)
        hy = _hash(3, y, seeds)
        assert np.array_equal(hx & mask, hy & mask)


def test_triple_hash_is_a_bijection_sample():
    x = np.arange(1 << 20, dtype=np.uint32) * np.uint32(2654435761)
    h = _hash(2, x)
    assert len(np.unique(h)) == x.size


def test_reverse_bits_and_zorder():
    x = np.array([1, 2, 0x80000000, 0x12345678], np.uint32)
    assert list(_hash(5, x)) == [0x80000000, 0x40000000, 1, 0x1E6A2C48]
    z = _hash(6, np.array([0b1011, 0xFF], np.uint32), np.array([0b0110, 0], np.uint32))
    assert list(z) == [0b01101101, 0x5555]
