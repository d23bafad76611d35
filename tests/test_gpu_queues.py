"""How the traversal kernels deal out their queues changes only which rays
share a wave.  A queue of at least 64 entries per wave of the grid is taken
in 64-entry batches by whichever wave asks next; a shorter one (the last
bounces' few paths) is spread as one batch of ceil(n / waves) entries per
wave.  With one traversal block per CU (512 waves on 256 CUs: queues of
32,768 rays or more are batched) the 128x72, 4 spp frame's first bounces
run batched and its later ones spread; with 64 blocks per CU every queue
is spread.  Frames rendered at 1, 64 and the default number of blocks per
CU are equal bit for bit, and equal to the oracle."""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu

W, H, SPP = 128, 72, 4
CH = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_PRIMITIVE | K.OSP_FB_ID_INSTANCE


def _scene():
    # all materials, the sun cone and four quad lights, instancing
    return scenes.c1(width=W, height=H, spp=SPP, grid=97, instances=40)


def _render(blocks_per_cu):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    if blocks_per_cu is not None:
        A.set_device_param(dev, "traceBlocksPerCU", blocks_per_cu)
        L.ospDeviceCommit(dev)
    try:
        A.check()
        sc = _scene()
        s = scenes.to_ospray(sc)
        fb = A.framebuffer(W, H, K.OSP_FB_RGBA32F, CH)
        A.render(fb, s.renderer, s.camera, s.world)
        return [A.map_color(fb, W, H, K.OSP_FB_RGBA32F)] + [A.map_channel(fb, c, W, H) for c in
                                                           (K.OSP_FB_DEPTH, K.OSP_FB_ID_PRIMITIVE,
                                                            K.OSP_FB_ID_INSTANCE)]
    finally:
        if blocks_per_cu is not None:
            A.set_device_param(dev, "traceBlocksPerCU", 12)
            L.ospDeviceCommit(dev)


def test_queue_batching_renders_the_same_frame(hip_device):
    base = _render(None)
    for bpc in (1, 64):
        other = _render(bpc)
        for k, (a, b) in enumerate(zip(base, other)):
            assert np.array_equal(a, b, equal_nan=True), (bpc, k)
    ofb, _, _ = orc.render_scene(_scene(), frames=1, fmt=K.OSP_FB_RGBA32F, channels=CH)
    assert np.array_equal(base[0], ofb.channel(K.OSP_FB_COLOR))
    for arr, ch in zip(base[1:], (K.OSP_FB_DEPTH, K.OSP_FB_ID_PRIMITIVE, K.OSP_FB_ID_INSTANCE)):
        assert np.array_equal(arr, ofb.channel(ch), equal_nan=True), ch
