"""The golden fixtures of tests/golden/ (written by make_golden.py from the
oracle) against the oracle and against what closed forms fix about them.

The oracle must still reproduce every fixture bit for bit (a change to the
oracle that moves any of these values shows up here first); the fixtures'
first hits agree with an FP64 brute force except on near-ties; the
low-discrepancy samples keep their stratification.  tests/test_gpu_golden.py
checks the device against the same files.
"""
import ctypes as C
import os

import numpy as np
import pytest

from ospray_amd import scenes
from oracle import orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


FRAME_CH = {"color": 1, "depth": 2, "normal": 16, "albedo": 32, "prim": 64, "obj": 128, "inst": 256,
            "first_normal": 512, "position": 1024}


def test_c0_hits_fixture():
    g = load("c0_hits.npz")
    osc = orc.OracleScene(scenes.c0(32, 32))
    tuv, ids, _ = osc.intersect(g["rays"])
    assert np.array_equal(tuv, g["tuv"]) and np.array_equal(ids, g["ids"])
    hit = g["fp64_ids"][:, 0] != 0xFFFFFFFF
    assert np.array_equal(hit, g["ids"][:, 0] != 0xFFFFFFFF)
    assert 0.2 < hit.mean() < 0.95
    t, t64 = g["tuv"][hit, 0].astype(np.float64), g["fp64_t"][hit].astype(np.float64)
    assert np.all(np.abs(t - t64) <= 1e-5 * t64)
    differ = ~(g["ids"] == g["fp64_ids"]).all(1)
    assert differ.sum() <= 2  # near-ties only


@pytest.mark.parametrize("name,scene,frames", [
    ("c0_frames.npz", lambda: scenes.c0(32, 32), 2),
    ("c1small_frames.npz", lambda: scenes.c1(width=48, height=27, spp=2, grid=97, instances=40), 2),
])
def test_frame_fixtures(name, scene, frames):
    g = load(name)
    ch = 4 | sum(FRAME_CH[k] for k in g if k in FRAME_CH) | (8 if "variance_estimate" in g else 0)
    fb, _, oren = orc.render_scene(scene(), frames=frames, fmt=3, channels=ch)
    for k, c in FRAME_CH.items():
        if k in g:
            assert np.array_equal(fb.channel(c), g[k], equal_nan=True), k
    if "variance_estimate" in g:
        assert np.float32(fb.variance(oren)) == g["variance_estimate"]
    col = g["color"]
    assert np.all(np.isfinite(col)) and col[..., 3].min() >= 0 and col[..., 3].max() <= 1
    # a pixel's ids are all invalid or all valid, and depth is infinite
    # exactly where nothing was hit
    miss = g["obj"] == 0xFFFFFFFF
    assert np.array_equal(miss, g["inst"] == 0xFFFFFFFF) and np.array_equal(miss, g["prim"] == 0xFFFFFFFF)
    assert np.array_equal(miss, np.isinf(g["depth"]))


def test_ld_sample_fixture():
    g = load("ld_samples.npz")
    buf = (C.c_float * 5)()
    for a, seed in enumerate(g["seeds"]):
        for i in range(g["samples"].shape[1]):
            orc.lib().orc_ld_samples5(int(seed), i, 16, buf)
            assert np.array_equal(np.frombuffer(buf, np.float32), g["samples"][a, i])
        # 64 scrambled Sobol points of one pixel: a (0,6,2)-net in dims 0/1
        s = g["samples"][a]
        for e in range(7):
            cells = np.floor(s[:, 0] * (1 << e)).astype(int) * (1 << (6 - e)) + np.floor(s[:, 1] * (1 << (6 - e))).astype(int)
            assert len(np.unique(cells)) == 64
