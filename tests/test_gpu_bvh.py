"""The device BVH build (hip_bvh.h: Morton sort, PLOC clustering, 4-wide
collapse) against the host binned-SAH build and the oracle.

Closest hits do not depend on the tree (ties go to the smallest primitive
key and the slab test is conservative), so the device-built BVH must give
the host-built BVH's hits and the oracle's bit for bit; its quality shows in
the node visits per ray, which the test bounds against the host build.
"""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu


def _host_bvh(on: bool):
    dev = K.lib().ospGetCurrentDevice()
    A.set_device_param(dev, "hostBvh", on)
    K.lib().ospDeviceCommit(dev)


def _random_rays(sc, n, seed, box=6.0):
    """Camera rays plus rays between random points of the scene box."""
    rng = np.random.default_rng(seed)
    cam = sc.camera
    pos = np.array(cam["position"], np.float64)
    d = np.array(cam["direction"], np.float64)
    d /= np.linalg.norm(d)
    du = np.cross(d, np.array(cam["up"], np.float64))
    du /= np.linalg.norm(du)
    dv = np.cross(du, d)
    h = np.tan(np.radians(cam["fovy"]) / 2)
    m = n // 2
    dirs = d + (rng.uniform(-1, 1, m) * h)[:, None] * du + (rng.uniform(-1, 1, m) * h)[:, None] * dv
    cam_rays = np.concatenate([np.tile(pos, (m, 1)), dirs / np.linalg.norm(dirs, axis=1, keepdims=True),
                               np.zeros((m, 1)), np.full((m, 1), np.inf)], 1)
    o = rng.uniform(-box, box, (n - m, 3))
    e = rng.normal(size=(n - m, 3))
    free = np.concatenate([o, e / np.linalg.norm(e, axis=1, keepdims=True), np.zeros((n - m, 1)),
                           rng.uniform(1, 30, (n - m, 1))], 1)
    return np.concatenate([cam_rays, free]).astype(np.float32)


def _mixed_scene():
    """One group holding a quad grid, an icosphere, 300 spheres and a
    3-triangle mesh: a BLAS over several geometries of every kind, plus a
    second group with the 3-triangle mesh alone (the smallest device build)
    and a third with 2 triangles (the host build takes those)."""
    rng = np.random.default_rng(3)
    n = 24
    xs = np.linspace(-6, 6, n + 1)
    gx, gz = np.meshgrid(xs, xs)
    gy = 0.3 * np.sin(gx) * np.cos(gz) - 5.0
    grid_v = np.stack([gx, gy, gz], -1).reshape(-1, 3).astype(np.float32)
    q = np.array([[j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i + 1, (j + 1) * (n + 1) + i]
                  for j in range(n) for i in range(n)], np.uint32)
    quads = scenes.Geometry("quads", vertex=grid_v, index=q)
    unit, faces = scenes.icosphere(3)
    ico = scenes.Geometry("triangles", vertex=(unit * 1.5 + [1.0, -1.0, 0.5]).astype(np.float32),
                          index=faces.astype(np.uint32))
    sph = np.concatenate([rng.uniform(-5, 5, (300, 3)), rng.uniform(0.05, 0.4, (300, 1))], 1).astype(np.float32)
    spheres = scenes.Geometry("spheres", spheres=sph)
    tri3 = scenes.Geometry("triangles", vertex=rng.uniform(-3, 3, (9, 3)).astype(np.float32),
                           index=np.arange(9, dtype=np.uint32).reshape(3, 3))
    tri2 = scenes.Geometry("triangles", vertex=rng.uniform(-3, 3, (6, 3)).astype(np.float32),
                           index=np.arange(6, dtype=np.uint32).reshape(2, 3))
    geoms = [quads, ico, spheres, tri3, tri2]
    mats = [{"type": "obj", "kd": (0.5, 0.5, 0.5)}]
    models = [scenes.Model(i, 0) for i in range(5)]
    groups = [[0, 1, 2, 3], [3], [4]]
    insts = [scenes.Instance(0), scenes.Instance(1, xfm=(1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 3, 0)),
             scenes.Instance(2, xfm=(2, 0, 0, 0, 2, 0, 0, 0, 2, -2, 2, 1))]
    cam = {"position": (0.0, 4.0, -20.0), "direction": (0.0, -0.2, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 45.0}
    return scenes.Scene("mixed", 64, 64, geoms, models, groups, insts, mats,
                        [{"type": "ambient", "intensity": 1.0}], cam, {"pixelSamples": 1, "maxPathLength": 2})


SCENES = {
    "c0": lambda: scenes.c0(32, 32),
    "c1_reduced": lambda: scenes.c1(width=96, height=54, spp=2, grid=193, instances=40),
    "mixed": _mixed_scene,
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_bvh_hits_equal_host_and_oracle(hip_device, name):
    sc = SCENES[name]()
    _host_bvh(False)
    sd = scenes.to_ospray(sc)
    ms, ndev, need = A.world_build(sd.world)
    _host_bvh(True)
    try:
        sh = scenes.to_ospray(sc)
        _, nhost, _ = A.world_build(sh.world)
    finally:
        _host_bvh(True)
    assert ndev >= 1 and nhost == 0
    assert 0 < need <= 128
    rays = _random_rays(sc, 20000, 11)
    tuv_d, ids_d, st_d = A.trace_rays(sd.world, rays)
    tuv_h, ids_h, st_h = A.trace_rays(sh.world, rays)
    assert np.array_equal(ids_d, ids_h) and np.array_equal(tuv_d, tuv_h)
    otuv, oids, _ = orc.OracleScene(sc).intersect(rays)
    assert np.array_equal(ids_d, oids) and np.array_equal(tuv_d, otuv)
    hit = ids_d[:, 0] != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.98
    # tree quality: node visits within 1.25x of the binned-SAH build
    assert st_d[0] <= 1.25 * st_h[0], (st_d, st_h)
    print(f"{name}: device BLAS build {ms:.2f} ms, node visits device/host {st_d[0] / st_h[0]:.3f}, "
          f"primitive tests {st_d[1] / st_h[1]:.3f}, stack need {need}")


def test_device_bvh_renders_like_host(hip_device):
    """A reduced 1080p scene over 2 frames: every channel identical with the
    device and the host BVH."""
    sc = scenes.c1(width=64, height=36, spp=2, grid=129, instances=24)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_PRIMITIVE | K.OSP_FB_ID_OBJECT
    out = []
    for host in (False, True):
        _host_bvh(host)
        try:
            s = scenes.to_ospray(sc)
            fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
            for _ in range(2):
                A.render(fb, s.renderer, s.camera, s.world)
            out.append([A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)]
                       + [A.map_channel(fb, c, sc.width, sc.height)
                          for c in (K.OSP_FB_DEPTH, K.OSP_FB_ID_PRIMITIVE, K.OSP_FB_ID_OBJECT)])
        finally:
            _host_bvh(True)
    for a, b in zip(*out):
        assert np.array_equal(a, b)
