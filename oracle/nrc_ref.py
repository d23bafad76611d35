"""ORACLE — test infrastructure only.

CPU reference of the neural radiance cache MLP (no counterpart in the
reference renderer; the design follows Mueller et al., "Real-time Neural
Radiance Caching for Path Tracing", SIGGRAPH 2021): frequency encoding of
position and direction, 4 hidden ReLU layers of 64, 3 linear outputs,
relative-L2 loss, and its backward pass, in FP64.  `bf16=True` rounds
weights, inputs, hidden activations and the back-propagated gradients to
bfloat16 (round to nearest even) where the device stores them in bf16, so
the device's fp32-accumulated results can be compared with a tolerance set
by summation order alone.  Pinned by central finite differences
(tests/test_nrc_cpu.py).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT, N_OUT_PAD, N_BANDS = 160, 64, 3, 16, 12
# parameter layout (elements): W0 [64][160], W1..W3 [64][64], W4 [16][64]
SHAPES = [(N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_OUT_PAD, N_HID)]
OFFSETS = np.cumsum([0] + [a * b for a, b in SHAPES]).tolist()
N_PARAMS = OFFSETS[-1]  # 23552
LOSS_EPS = 0.01


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray) -> list:
    return [params[OFFSETS[i]:OFFSETS[i + 1]].reshape(SHAPES[i]) for i in range(5)]


def encode(raw: np.ndarray, lo, hi) -> np.ndarray:
    """raw: (n, 12) = position, direction, normal, albedo.  Returns (n, 160):
    for each of the 6 coordinates (position mapped to [0, 1] over the scene
    box [lo, hi], then direction) sin and cos of 2^k pi x for k < 12, then
    albedo, normal, the mapped position, direction, a constant 1, 3 zeros."""
    raw = np.asarray(raw, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    p = (raw[:, 0:3] - lo) / (hi - lo)
    d = raw[:, 3:6]
    coords = np.concatenate([p, d], 1)  # (n, 6)
    n = raw.shape[0]
    out = np.zeros((n, N_IN))
    for c in range(6):
        for k in range(N_BANDS):
            a = (2.0 ** k) * np.pi * coords[:, c]
            out[:, c * 2 * N_BANDS + 2 * k] = np.sin(a)
            out[:, c * 2 * N_BANDS + 2 * k + 1] = np.cos(a)
    out[:, 144:147] = raw[:, 9:12]
    out[:, 147:150] = raw[:, 6:9]
    out[:, 150:153] = p
    out[:, 153:156] = d
    out[:, 156] = 1.0
    return out


def forward(params, x, bf16=False):
    """Returns (y (n, 3), hidden activations [h1..h4])."""
    W = unpack(np.asarray(params, np.float64))
    r = bf16_round if bf16 else (lambda a: a)
    W = [r(w) for w in W]
    h = r(np.asarray(x, np.float64))
    hs = []
    for i in range(4):
        h = r(np.maximum(h @ W[i].T, 0.0))
        hs.append(h)
    y = h @ W[4].T
    return y[:, :N_OUT], hs


def loss_and_grad(params, x, target, bf16=False):
    """Relative L2 loss mean over batch and channels,
    L = mean((y - t)^2 / (y^2 + eps)) with the denominator held constant in
    the gradient; returns (L, dL/dparams) in the flat layout."""
    W = unpack(np.asarray(params, np.float64))
    r = bf16_round if bf16 else (lambda a: a)
    Wr = [r(w) for w in W]
    x = r(np.asarray(x, np.float64))
    n = x.shape[0]
    y, hs = forward(params, x, bf16)
    t = np.asarray(target, np.float64)
    den = y * y + LOSS_EPS
    scale = 1.0 / (N_OUT * n)
    L = float(np.sum((y - t) ** 2 / den) * scale)
    dy = np.zeros((n, N_OUT_PAD))
    dy[:, :N_OUT] = r(2.0 * (y - t) / den * scale)
    grads = [None] * 5
    acts = [x] + hs  # inputs of W0..W4
    dz = dy
    for i in range(4, -1, -1):
        grads[i] = dz.T @ acts[i]
        if i > 0:
            dh = dz @ Wr[i]
            dz = r(dh * (acts[i] > 0))
    return L, np.concatenate([g.ravel() for g in grads])


def loss_value(params, x, target) -> float:
    y, _ = forward(params, x)
    t = np.asarray(target, np.float64)
    return float(np.mean((y - t) ** 2 / (y * y + LOSS_EPS)))


def init_params(seed: int = 1) -> np.ndarray:
    """He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)); padded output
    rows zero.  (The device draws its own initial weights; this is for
    CPU-side tests.)"""
    rng = np.random.default_rng(seed)
    out = []
    for i, (a, b) in enumerate(SHAPES):
        lim = np.sqrt(6.0 / b)
        w = rng.uniform(-lim, lim, (a, b))
        if i == 4:
            w[N_OUT:] = 0.0
        out.append(w.ravel())
    return np.concatenate(out)
