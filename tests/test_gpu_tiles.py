"""The pixel tile of the path slots (device parameter tileWidth: 64-slot
tiles of 4x16, 8x8, 16x4 or 32x2 pixels) changes only which rays share a
wave: frames rendered with each tile are equal bit for bit, including a
frame whose size is no multiple of any tile (the padded slots of the last
tile row and column render nothing), and equal to the oracle."""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu

W, H = 37, 23


def _render(tile_width):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "tileWidth", tile_width)
    L.ospDeviceCommit(dev)
    try:
        sc = scenes.c0(W, H)
        s = scenes.to_ospray(sc, {"pixelSamples": 2, "maxPathLength": 3})
        ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_PRIMITIVE
        fb = A.framebuffer(W, H, K.OSP_FB_RGBA32F, ch)
        for _ in range(2):
            A.render(fb, s.renderer, s.camera, s.world)
        return (A.map_color(fb, W, H, K.OSP_FB_RGBA32F), A.map_channel(fb, K.OSP_FB_DEPTH, W, H),
                A.map_channel(fb, K.OSP_FB_ID_PRIMITIVE, W, H)), sc, ch
    finally:
        A.set_device_param(dev, "tileWidth", 8)
        L.ospDeviceCommit(dev)


def test_tile_widths_render_the_same_frame(hip_device):
    base, sc, ch = _render(8)
    for tw in (4, 16, 32):
        other, _, _ = _render(tw)
        for a, b in zip(base, other):
            assert np.array_equal(a, b), tw
    ofb, _, _ = orc.render_scene(sc, frames=2, fmt=K.OSP_FB_RGBA32F, channels=ch,
                                 renderer_params={"pixelSamples": 2, "maxPathLength": 3})
    assert np.array_equal(base[2], ofb.channel(K.OSP_FB_ID_PRIMITIVE))
    assert np.array_equal(base[1], ofb.channel(K.OSP_FB_DEPTH))
    assert float(np.abs(base[0] - ofb.channel(K.OSP_FB_COLOR)).max()) <= 1e-5


def test_invalid_tile_width_is_an_error(hip_device):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "tileWidth", 12)
    L.ospDeviceCommit(dev)
    with pytest.raises(A.OSPRayError, match="tileWidth"):
        A.check()
    A.set_device_param(dev, "tileWidth", 8)
    L.ospDeviceCommit(dev)
    A.check()
