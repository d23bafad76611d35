"""The device against the committed golden fixtures (tests/golden/), through
the C API, without running the oracle: first hits of the fixture's rays and
the accumulated frame-buffer channels, bit for bit."""
import os

import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FRAME_CH = {"color": 1, "depth": 2, "normal": 16, "albedo": 32, "prim": 64, "obj": 128, "inst": 256,
            "first_normal": 512, "position": 1024}


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_c0_hits_golden(hip_device):
    g = load("c0_hits.npz")
    s = scenes.to_ospray(scenes.c0(32, 32))
    tuv, ids, _ = A.trace_rays(s.world, g["rays"])
    assert np.array_equal(ids, g["ids"])
    assert np.array_equal(tuv, g["tuv"])


@pytest.mark.parametrize("name,scene,frames", [
    ("c0_frames.npz", lambda: scenes.c0(32, 32), 2),
    ("c1small_frames.npz", lambda: scenes.c1(width=48, height=27, spp=2, grid=97, instances=40), 2),
])
def test_frames_golden(hip_device, name, scene, frames):
    g = load(name)
    sc = scene()
    ch = 4 | sum(FRAME_CH[k] for k in g if k in FRAME_CH) | (8 if "variance_estimate" in g else 0)
    s = scenes.to_ospray(sc)
    fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    for _ in range(frames):
        A.render(fb, s.renderer, s.camera, s.world)
    assert np.array_equal(A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F), g["color"])
    for k, c in FRAME_CH.items():
        if k in g and k != "color":
            assert np.array_equal(A.map_channel(fb, c, sc.width, sc.height), g[k], equal_nan=True), k
    if "variance_estimate" in g:
        assert np.float32(K.lib().ospGetVariance(fb.h)) == g["variance_estimate"]
