"""Near-tie re-resolution on the GPU: closest hits of extension rays whose
float distances tie within 1e-5 are re-resolved by FP64 distances (the
second traversal of pt_trace.h), bit for bit as the oracle does, and equal
to the FP64 brute force on surfaces one ulp apart."""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

from test_oracle_neartie import rays_towards_layers

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("instanced", [False, True])
def test_near_tie_first_hits(hip_device, instanced):
    sc = scenes.near_tie_layers(instanced=instanced)
    s = scenes.to_ospray(sc)
    osc = orc.OracleScene(sc)
    rays = rays_towards_layers(4000, 3)
    tuv, ids, _ = A.trace_rays(s.world, rays)
    otuv, oids, _ = osc.intersect(rays)
    _, fids, _ = osc.intersect(rays, fp64=True)
    assert np.array_equal(ids, oids) and np.array_equal(tuv, otuv)
    assert np.array_equal(ids, fids)


def test_near_tie_render(hip_device):
    sc = scenes.near_tie_layers(instanced=True)
    ch = (K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_PRIMITIVE | K.OSP_FB_ID_OBJECT
          | K.OSP_FB_ID_INSTANCE | K.OSP_FB_NORMAL | K.OSP_FB_FIRST_NORMAL)
    s = scenes.to_ospray(sc)
    fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    for _ in range(2):
        A.render(fb, s.renderer, s.camera, s.world)
    ofb, _, _ = orc.render_scene(sc, frames=2, fmt=K.OSP_FB_RGBA32F, channels=ch)
    assert np.array_equal(A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F), ofb.channel(K.OSP_FB_COLOR))
    for c in (K.OSP_FB_DEPTH, K.OSP_FB_ID_PRIMITIVE, K.OSP_FB_ID_OBJECT, K.OSP_FB_ID_INSTANCE, K.OSP_FB_NORMAL,
              K.OSP_FB_FIRST_NORMAL):
        assert np.array_equal(A.map_channel(fb, c, sc.width, sc.height), ofb.channel(c)), c
    # every visible layer pixel shows the front layer or the outer sphere
    obj = ofb.channel(K.OSP_FB_ID_OBJECT)
    assert set(np.unique(obj[obj != 0xFFFFFFFF])) <= {3, 4}
