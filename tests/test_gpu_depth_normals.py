"""First-hit depth and normals of the GPU frame buffer against FP64.

C0 (64 icospheres and a ground quad; the icospheres' vertex normals
dropped, so every shading normal is the geometric one) at 96x96 with the
point pixel filter, 1 spp, one frame: pixel (x, y) is the camera ray
through the screen position ((x + 0.5) / 96, (y + 0.5) / 96), formed by the
oracle's camera in float exactly as on the device.  The FP64 brute force
over every primitive gives the hit primitive and distance; the normal is
the FP64 cross product of the primitive's edges, normalised and turned
against the ray (postIntersect, common/World.ih:389-465).  The frame
buffer's primitive and model ids equal the FP64 ones, depth agrees within
1e-5 relative and the first-hit normal within 1e-5 per component.
"""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu

N = 96


def test_first_hit_depth_and_normals_against_fp64(hip_device):
    sc = scenes.c0(N, N)
    for g in sc.geometries:
        g.normal = None
    rp = {"pixelSamples": 1, "maxPathLength": 1, "pixelFilter": A.UInt(K.OSP_PIXELFILTER_POINT)}
    s = scenes.to_ospray(sc, rp)
    ch = (K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_FIRST_NORMAL | K.OSP_FB_ID_PRIMITIVE
          | K.OSP_FB_ID_OBJECT)
    fb = A.framebuffer(N, N, K.OSP_FB_RGBA32F, ch)
    A.render(fb, s.renderer, s.camera, s.world)
    depth = A.map_channel(fb, K.OSP_FB_DEPTH, N, N).reshape(-1)
    nrm = A.map_channel(fb, K.OSP_FB_FIRST_NORMAL, N, N).reshape(-1, 3).astype(np.float64)
    prim = A.map_channel(fb, K.OSP_FB_ID_PRIMITIVE, N, N).reshape(-1)
    obj = A.map_channel(fb, K.OSP_FB_ID_OBJECT, N, N).reshape(-1)

    ys, xs = np.mgrid[0:N, 0:N]
    screen = np.stack([(xs.reshape(-1) + 0.5) / N, (ys.reshape(-1) + 0.5) / N], 1).astype(np.float32)
    cam = orc.camera_rays(sc.camera, screen, near_clip=1e-6)
    rays = np.concatenate([cam, np.full((N * N, 1), 1e-6, np.float32), np.full((N * N, 1), np.inf, np.float32)], 1)
    tuv64, ids64, _ = orc.OracleScene(sc).intersect(rays, fp64=True)
    hit = ids64[:, 0] != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.99
    assert np.array_equal(prim[hit], ids64[hit, 2]) and np.array_equal(obj[hit], ids64[hit, 1])
    assert np.all(prim[~hit] == 0xFFFFFFFF)
    t64 = tuv64[hit, 0].astype(np.float64)
    assert np.abs(depth[hit] - t64).max() <= 1e-5 * t64.max()
    assert np.all(np.abs(depth[hit] - t64) <= 1e-5 * t64)

    # FP64 geometric normals, turned against the ray
    idx = np.where(hit)[0]
    n64 = np.empty((len(idx), 3))
    for j, i in enumerate(idx):
        g = sc.geometries[sc.models[int(ids64[i, 1])].geometry]
        v = g.vertex.astype(np.float64)[g.index[int(ids64[i, 2])]]
        if len(v) == 4:  # a planar quad: the plane of (v0, v1, v3)
            v = v[[0, 1, 3]]
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n /= np.linalg.norm(n)
        if np.dot(rays[i, 3:6].astype(np.float64), n) >= 0.0:
            n = -n
        n64[j] = n
    err = np.abs(nrm[idx] - n64).max()
    print(f"first hits of {len(idx)} pixels: depth max rel err "
          f"{(np.abs(depth[hit] - t64) / t64).max():.2e}, normal max abs err {err:.2e}")
    assert err <= 1e-5
