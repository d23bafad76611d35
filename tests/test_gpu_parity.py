"""GPU parity: the hip device against the CPU oracle, through the C API.

Bar: first-hit ids, depth, normals and every framebuffer channel equal to
the oracle bit for bit on the same seeded inputs (the device's arithmetic is
written to produce the oracle's float results; see pt_math.h).  Full-size
workloads are checked through first-hit parity on ray subsets and through
size-independent properties.
"""
import os
import subprocess

import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes
from oracle import orc

pytestmark = pytest.mark.gpu

ALL_CH = (K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_VARIANCE | K.OSP_FB_NORMAL | K.OSP_FB_ALBEDO
          | K.OSP_FB_ID_PRIMITIVE | K.OSP_FB_ID_OBJECT | K.OSP_FB_ID_INSTANCE | K.OSP_FB_FIRST_NORMAL
          | K.OSP_FB_POSITION)
AOV = [K.OSP_FB_DEPTH, K.OSP_FB_NORMAL, K.OSP_FB_ALBEDO, K.OSP_FB_ID_PRIMITIVE, K.OSP_FB_ID_OBJECT,
       K.OSP_FB_ID_INSTANCE, K.OSP_FB_FIRST_NORMAL, K.OSP_FB_POSITION]


def _camera_rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    cam = sc.camera
    pos = np.array(cam["position"], np.float64)
    d = np.array(cam["direction"], np.float64)
    d /= np.linalg.norm(d)
    up = np.array(cam["up"], np.float64)
    du = np.cross(d, up)
    du /= np.linalg.norm(du)
    dv = np.cross(du, d)
    h = np.tan(np.radians(cam["fovy"]) / 2)
    a = cam.get("aspect", 1.0)
    sx, sy = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    dirs = d + (sx * h * a)[:, None] * du + (sy * h)[:, None] * dv
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return np.concatenate([np.tile(pos, (n, 1)), dirs, np.zeros((n, 1)), np.full((n, 1), np.inf)], 1).astype(np.float32)


def _render_both(sc, frames, fmt=K.OSP_FB_RGBA32F, channels=ALL_CH, renderer_params=None):
    s = scenes.to_ospray(sc, renderer_params)
    fb = A.framebuffer(sc.width, sc.height, fmt, channels)
    for _ in range(frames):
        A.render(fb, s.renderer, s.camera, s.world)
    ofb, osc, oren = orc.render_scene(sc, frames=frames, fmt=fmt, channels=channels, renderer_params=renderer_params)
    return s, fb, ofb, osc, oren


def _assert_channels_equal(fb, ofb, sc, channels):
    w, h = sc.width, sc.height
    for ch in AOV:
        if channels & ch:
            g = A.map_channel(fb, ch, w, h)
            o = ofb.channel(ch)
            assert np.array_equal(g, o, equal_nan=True), f"channel {ch}: {np.sum(g != o)} pixels differ"


@pytest.mark.parametrize("seed", [1, 2])
def test_first_hit_c0(hip_device, seed):
    sc = scenes.c0(32, 32)
    s = scenes.to_ospray(sc)
    osc = orc.OracleScene(sc)
    rays = _camera_rays(sc, 20000, seed)
    tuv, ids, stats = A.trace_rays(s.world, rays)
    otuv, oids, _ = osc.intersect(rays)
    assert np.array_equal(ids, oids)
    assert np.array_equal(tuv, otuv)
    assert stats[0] > 0 and stats[1] > 0
    # against the FP64 brute force: same ids except near-ties
    sub = rays[:2000]
    ftuv, fids, _ = osc.intersect(sub, fp64=True)
    same = (ids[:2000] == fids).all(axis=1)
    hit = fids[:, 0] != 0xFFFFFFFF
    assert np.all(np.abs(tuv[:2000][hit, 0] - ftuv[hit, 0]) <= 1e-5 * ftuv[hit, 0])
    for i in np.where(~same)[0]:
        assert abs(tuv[i, 0] - ftuv[i, 0]) <= 1e-5 * ftuv[i, 0]


def test_render_c0_bit_exact(hip_device):
    sc = scenes.c0(64, 64)
    s, fb, ofb, _, oren = _render_both(sc, frames=4)
    col = A.map_color(fb, 64, 64, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert K.lib().ospGetVariance(fb.h) == pytest.approx(ofb.variance(oren), rel=1e-6)
    assert np.isfinite(K.lib().ospGetVariance(fb.h))


def test_render_c1_small_bit_exact(hip_device):
    # all materials (Lambert, smooth and rough metal, glass), the sun cone and
    # four quad lights, instancing
    sc = scenes.c1(width=96, height=54, spp=2, grid=97, instances=40)
    s, fb, ofb, _, _ = _render_both(sc, frames=3)
    col = A.map_color(fb, 96, 54, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)


@pytest.mark.parametrize("fmt", [K.OSP_FB_RGBA8, K.OSP_FB_SRGBA])
def test_color_formats(hip_device, fmt):
    sc = scenes.c0(40, 24)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM
    s, fb, ofb, _, _ = _render_both(sc, frames=2, fmt=fmt, channels=ch)
    assert np.array_equal(A.map_color(fb, 40, 24, fmt), ofb.channel(K.OSP_FB_COLOR))


def test_renderer_parameters(hip_device):
    # light sample count, max contribution, background colour, box filter,
    # projected depth off, no accumulation (blue-noise sample sequence)
    sc = scenes.c0(32, 32, spp=3)
    rp = {"lightSamples": 3, "maxContribution": 2.0, "backgroundColor": (0.1, 0.2, 0.3, 1.0),
          "pixelFilter": A.UInt(K.OSP_PIXELFILTER_BOX), "roulettePathLength": 1, "minContribution": 0.01}
    ch = K.OSP_FB_COLOR | K.OSP_FB_DEPTH | K.OSP_FB_ID_OBJECT
    s, fb, ofb, _, _ = _render_both(sc, frames=1, channels=ch, renderer_params=rp)
    assert np.array_equal(A.map_color(fb, 32, 32, K.OSP_FB_RGBA32F), ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ch)


def test_reset_accumulation(hip_device):
    sc = scenes.c0(24, 24)
    s = scenes.to_ospray(sc)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM
    fb = A.framebuffer(24, 24, K.OSP_FB_RGBA32F, ch)
    for _ in range(3):
        A.render(fb, s.renderer, s.camera, s.world)
    K.lib().ospResetAccumulation(fb.h)
    A.render(fb, s.renderer, s.camera, s.world)
    ofb, _, _ = orc.render_scene(sc, frames=1, fmt=K.OSP_FB_RGBA32F, channels=ch)
    assert np.array_equal(A.map_color(fb, 24, 24, K.OSP_FB_RGBA32F), ofb.channel(K.OSP_FB_COLOR))


def test_adaptive_accumulation(hip_device):
    # varianceThreshold between the task errors: from frame 2 on (spp 8)
    # only the noisier 8x8 tasks are rendered
    sc = scenes.c0(48, 40, spp=8)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_VARIANCE | K.OSP_FB_DEPTH | K.OSP_FB_ID_OBJECT
    ofb, _, _ = orc.render_scene(sc, frames=2, channels=ch)
    thr = float(np.median(ofb.task_variance()))
    rp = {"varianceThreshold": thr}
    s, fb, ofb, _, oren = _render_both(sc, frames=5, channels=ch, renderer_params=rp)
    assert np.array_equal(A.map_color(fb, 48, 40, K.OSP_FB_RGBA32F), ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ch)
    assert K.lib().ospGetVariance(fb.h) == pytest.approx(ofb.variance(oren), rel=1e-6)


def _c0_with_emitters(w=48, h=48, spp=2, max_path_length=3):
    """C0 with three luminous icospheres (one semi-transparent), a luminous
    sphere primitive, and a second, scaled instance of the whole group: six
    geometry lights before the quad and ambient lights; a model instanced
    twice takes the area pdf of its last instance."""
    sc = scenes.c0(w, h, spp=spp, max_path_length=max_path_length)
    for k, t in ((3, 0.0), (17, 0.3), (40, 0.0)):
        sc.materials[sc.models[k].material] = {"type": "luminous", "intensity": 3.0 + k * 0.1,
                                               "color": (1.0, 0.8, 0.6), "transparency": t}
    sc.geometries.append(scenes.Geometry("spheres", spheres=np.array([[2.0, 5.0, -3.0, 0.7]], np.float32)))
    sc.materials.append({"type": "luminous", "intensity": 2.0})
    sc.models.append(scenes.Model(geometry=len(sc.geometries) - 1, material=len(sc.materials) - 1))
    sc.groups.append([len(sc.models) - 1])
    sc.instances.append(scenes.Instance(1))
    sc.instances.append(scenes.Instance(0, xfm=(0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 6.0, 1.0, 4.0)))
    return sc


@pytest.mark.parametrize("geometry_lights", [True, False])
def test_geometry_lights_bit_exact(hip_device, geometry_lights):
    sc = _c0_with_emitters()
    rp = {"geometryLights": geometry_lights}
    s, fb, ofb, _, oren = _render_both(sc, frames=3, renderer_params=rp)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert col[..., :3].max() > 1.0  # emitters seen directly


def _c0_transparent_layers(w=48, h=48, spp=2, max_path_length=3):
    """C0 under transparent occluders of the quad light's shadow rays: six
    stacked obj quads (d 0.8), and in a second instance (scaled in x and z)
    two more layers and two transparent sphere primitives, a large one
    around part of the ground and a small one above the layers.  Shadow
    rays from the ground cross more hits than the occlusion pass records,
    so the transparency pass traverses again; those from the upper layers
    and the small sphere cross four or fewer, which the occlusion pass
    records (rays leaving the small sphere's lit side start inside it:
    exit root only)."""
    sc = scenes.c0(w, h, spp=spp, max_path_length=max_path_length)
    def layers(n, y0):
        v, ix = [], []
        for k in range(n):
            y, b = y0 + k, 4 * k
            v += [[-6, y, -6], [6, y, -6], [6, y, 6], [-6, y, 6]]
            ix.append([b, b + 1, b + 2, b + 3])
        return scenes.Geometry("quads", vertex=np.array(v, np.float32), index=np.array(ix, np.uint32))
    sc.materials.append({"type": "obj", "kd": (0.6, 0.7, 0.8), "d": 0.8, "tf": (0.3, 0.2, 0.1)})
    mat = len(sc.materials) - 1
    sc.materials.append({"type": "obj", "kd": (0.5, 0.5, 0.3), "d": 0.6, "tf": (0.4, 0.4, 0.4)})
    smat = len(sc.materials) - 1
    sc.geometries += [layers(6, 1.0), layers(2, 1.5),
                      scenes.Geometry("spheres", spheres=np.array([[0.0, -6.0, 0.0, 3.0], [0.0, 7.0, 0.0, 0.6]],
                                                                   np.float32))]
    g = len(sc.geometries)
    sc.models += [scenes.Model(geometry=g - 3, material=mat), scenes.Model(geometry=g - 2, material=mat),
                  scenes.Model(geometry=g - 1, material=smat)]
    m = len(sc.models)
    sc.groups += [[m - 3], [m - 2, m - 1]]
    sc.instances += [scenes.Instance(len(sc.groups) - 2),
                     scenes.Instance(len(sc.groups) - 1, xfm=(0.5, 0, 0, 0, 1, 0, 0, 0, 0.5, 1.0, 0.0, -1.0))]
    return sc


@pytest.mark.parametrize("max_path_length", [3, 8])
def test_transparent_shadows_bit_exact(hip_device, max_path_length):
    # recorded hits (up to 4 per shadow ray) and the traversal fallback of
    # the transparency pass, against the oracle's closest-hit loop
    sc = _c0_transparent_layers(max_path_length=max_path_length)
    s, fb, ofb, _, oren = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert col[..., :3].mean() > 0.01


def _edge_scene(name):
    import dataclasses
    if name == "empty_world":  # lights only: the background and the ambient light
        sc = scenes.c0(16, 12, spp=2, max_path_length=3)
        return dataclasses.replace(sc, geometries=[], models=[], groups=[], instances=[], materials=[])
    if name == "no_lights":
        return dataclasses.replace(scenes.c0(16, 12, spp=2, max_path_length=3), lights=[])
    if name == "1x1":
        return scenes.c0(1, 1, spp=3, max_path_length=4)
    if name == "3x2":
        return scenes.c0(3, 2, spp=2, max_path_length=4)
    if name == "depth0":
        return scenes.c0(16, 12, spp=2, max_path_length=0)
    return scenes.c0(16, 12, spp=2, max_path_length=1)


@pytest.mark.parametrize("name", ["empty_world", "no_lights", "1x1", "3x2", "depth0", "depth1"])
def test_edge_cases_bit_exact(hip_device, name):
    # an empty world, a world without lights, framebuffers of one pixel and
    # of less than one 8x8 task, and paths of length 0 and 1: every channel
    # and the variance equal the oracle's
    sc = _edge_scene(name)
    s, fb, ofb, _, oren = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert K.lib().ospGetVariance(fb.h) == pytest.approx(ofb.variance(oren), rel=1e-6, nan_ok=True)


def _c0_specular(max_path_length):
    # the C0 spheres with OBJ specular lobes: ks, ns, partial opacity d and
    # transmission tf, and one material whose kd + ks exceeds 1 (the energy
    # normalisation of OBJ.cpp:39-47)
    sc = scenes.c0(24, 16, spp=2, max_path_length=max_path_length)
    for i, m in enumerate(sc.materials[:-1]):
        k = i % 4
        if k == 0:
            m.update(ks=(0.3, 0.3, 0.3), ns=10.0)
        elif k == 1:
            m.update(ks=(0.5, 0.4, 0.2), ns=50.0, d=0.7, tf=(0.2, 0.3, 0.4))
        elif k == 2:
            m.update(kd=(0.9, 0.8, 0.7), ks=(0.6, 0.6, 0.6), ns=3.0)
    sc.materials[-1].update(ks=(0.2, 0.2, 0.2), ns=100.0)
    return sc


@pytest.mark.parametrize("max_path_length", [2, 6])
def test_obj_specular_bit_exact(hip_device, max_path_length):
    # the MultiBSDF with the Blinn-Phong child: sampled by choice, evaluated
    # in the sample's combination and in next-event estimation
    sc = _c0_specular(max_path_length)
    s, fb, ofb, _, oren = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert K.lib().ospGetVariance(fb.h) == pytest.approx(ofb.variance(oren), rel=1e-6, nan_ok=True)
    assert col[..., :3].mean() > 0.01


# a copper-like (wavelength nm, eta, k) table, ordered by wavelength
_COPPER_IOR = np.array([[400.0, 1.18, 2.21], [450.0, 1.17, 2.40], [500.0, 1.13, 2.57], [550.0, 0.72, 2.62],
                        [600.0, 0.27, 3.24], [650.0, 0.21, 3.67], [700.0, 0.21, 4.05]], np.float32)


@pytest.mark.parametrize("max_path_length", [2, 5])
def test_spectral_metal_bit_exact(hip_device, max_path_length):
    # metal with an 'ior' table (Metal.cpp:34-57): smooth and rough, beside
    # RGB metals and Lambertian spheres
    sc = scenes.c0(24, 16, spp=2, max_path_length=max_path_length)
    for i in range(0, 64, 3):
        sc.materials[i] = {"type": "metal", "roughness": (0.0, 0.25, 0.6)[(i // 3) % 3],
                           "ior": _COPPER_IOR}
    for i in range(1, 64, 6):
        sc.materials[i] = {"type": "metal", "roughness": 0.3}
    s, fb, ofb, _, oren = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert col[..., :3].mean() > 0.01


def _set_device_params(**kv):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    for k, v in kv.items():
        A.set_device_param(dev, k, v)
    L.ospDeviceCommit(dev)


@pytest.mark.parametrize("lanes,budget_mb", [(2, 16384), (3, 16384), (2, 1)])
def test_pass_pipelining_bit_exact(hip_device, lanes, budget_mb):
    # a frame's passes alternate over lanes (the accumulation stays in
    # sample order): every channel equals the one-lane frames and the
    # oracle; a 1 MB path budget gives one sample per pass, more passes
    # than lanes
    sc = scenes.c0(40, 36, spp=5, max_path_length=4)
    images = {}
    try:
        for fl in (1, lanes):
            _set_device_params(frameLanes=fl, pathBudgetMB=budget_mb)
            s = scenes.to_ospray(sc)
            fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ALL_CH)
            for _ in range(2):
                A.render(fb, s.renderer, s.camera, s.world)
            images[fl] = [A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)] + \
                [A.map_channel(fb, ch, sc.width, sc.height) for ch in AOV] + [K.lib().ospGetVariance(fb.h)]
    finally:
        _set_device_params(frameLanes=1, pathBudgetMB=16384)
    for a, b in zip(images[1], images[lanes]):
        assert np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)
    ofb, _, _ = orc.render_scene(sc, frames=2, fmt=K.OSP_FB_RGBA32F, channels=ALL_CH)
    assert np.array_equal(images[lanes][0], ofb.channel(K.OSP_FB_COLOR))


def _volume(anisotropy=0.3, user_id=None, seed=3):
    """A 24x22x20 random grid (box smoothed) with an 8-point transfer
    function: non-cubic, so a swapped axis shows."""
    rng = np.random.default_rng(seed)
    d = rng.random((20, 22, 24)).astype(np.float32)
    d = (d + np.roll(d, 1, 0) + np.roll(d, 1, 1) + np.roll(d, 1, 2)) * 0.25
    col = rng.uniform(0.2, 1.0, (8, 3)).astype(np.float32)
    op = (rng.random(8) * 0.5).astype(np.float32)
    return scenes.Volume(d, col, op, origin=(-1.0, -1.0, -1.0), spacing=(0.1, 0.09, 0.11), tf_range=(0.15, 0.85),
                         density_scale=1.5, anisotropy=anisotropy, user_id=user_id)


def _add_volume_instance(sc, vol, xfm, inst_id=None):
    if vol not in sc.volumes:
        sc.volumes.append(vol)
        sc.groups.append([])
        sc.group_volumes = sc.group_volumes + [[] for _ in range(len(sc.groups) - 1 - len(sc.group_volumes))]
        sc.group_volumes.append([len(sc.volumes) - 1])
    gi = len(sc.groups) - 1
    sc.instances.append(scenes.Instance(gi, xfm=xfm, user_id=inst_id))


@pytest.mark.parametrize("anisotropy", [0.3, -0.6, 0.0])
def test_volume_in_c0_bit_exact(hip_device, anisotropy):
    """C0 with a scaled, translated heterogeneous volume around some of the
    spheres: free paths, HG scattering, NEE and virtual-light transmittance."""
    sc = scenes.c0(48, 48, spp=2, max_path_length=5)
    _add_volume_instance(sc, _volume(anisotropy), (4, 0, 0, 0, 4, 0, 0, 0, 4, -1.0, -2.0, 0.0))
    s, fb, ofb, _, _ = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    # some camera rays collide in the volume (model id = group-local 0 of the
    # volume's instance)
    inst = A.map_channel(fb, K.OSP_FB_ID_INSTANCE, sc.width, sc.height)
    assert (inst == len(sc.instances) - 1).sum() > 20


def test_two_volume_instances_with_ids_bit_exact(hip_device):
    """The same volumetric model instanced twice (overlapping boxes, one
    rotated), with model and instance ids."""
    sc = scenes.c0(40, 40, spp=2, max_path_length=4)
    vol = _volume(0.5, user_id=7, seed=4)
    _add_volume_instance(sc, vol, (4, 0, 0, 0, 4, 0, 0, 0, 4, -1.0, -2.0, 0.0), inst_id=21)
    c, sn = 0.8, 0.6
    _add_volume_instance(sc, vol, (3 * c, 3 * sn, 0, -3 * sn, 3 * c, 0, 0, 0, 3, 1.0, 0.0, -2.0))
    s, fb, ofb, _, _ = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    obj = A.map_channel(fb, K.OSP_FB_ID_OBJECT, sc.width, sc.height)
    assert (obj == 7).sum() > 10


def test_volume_only_world_bit_exact(hip_device):
    """A world with no geometry: one volume lit by a quad and an ambient light."""
    sc = scenes.c0(32, 32, spp=4, max_path_length=6)
    sc.geometries, sc.models, sc.materials = [], [], []
    sc.groups, sc.instances, sc.group_volumes = [], [], []
    _add_volume_instance(sc, _volume(0.0, seed=5), (5, 0, 0, 0, 5, 0, 0, 0, 5, -2.0, -3.0, -2.0))
    s, fb, ofb, _, _ = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)
    assert col[..., :3].max() > 0.0


def test_c4_small_bit_exact(hip_device):
    """The surface+volume configuration at reduced size: 64^3 fBm grid,
    96x54, 2 spp, maxPathLength 10."""
    sc = scenes.c4(96, 54, spp=2, grid=64, base_period=16)
    s, fb, ofb, _, _ = _render_both(sc, frames=2)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.array_equal(col, ofb.channel(K.OSP_FB_COLOR))
    _assert_channels_equal(fb, ofb, sc, ALL_CH)


def test_c4_full_size_band_bit_exact(hip_device):
    """The full surface+volume configuration (512^3 grid, 1280x720, 2 spp):
    the GPU frame's rows 352-367 against the oracle's."""
    sc = scenes.c4()
    s = scenes.to_ospray(sc)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_OBJECT | K.OSP_FB_ID_INSTANCE
    fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    A.render(fb, s.renderer, s.camera, s.world)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    assert np.isfinite(col).all()
    rows = (352, 368)
    ofb, _, _ = orc.render_scene(sc, frames=1, channels=ch, rows=rows)
    o = ofb.channel(K.OSP_FB_COLOR)
    assert np.array_equal(col[rows[0]:rows[1]], o[rows[0]:rows[1]])
    for c in (K.OSP_FB_DEPTH, K.OSP_FB_ID_OBJECT, K.OSP_FB_ID_INSTANCE):
        g = A.map_channel(fb, c, sc.width, sc.height)[rows[0]:rows[1]]
        assert np.array_equal(g, ofb.channel(c)[rows[0]:rows[1]], equal_nan=True), c
    assert (A.map_channel(fb, K.OSP_FB_ID_INSTANCE, sc.width, sc.height) == 1).mean() > 0.05


def test_pick(hip_device):
    sc = scenes.c0(64, 64)
    s = scenes.to_ospray(sc)
    fb = A.framebuffer(64, 64)
    res = K.OSPPickResult()
    K.lib().ospPick(res, fb.h, s.renderer.h, s.camera.h, s.world.h, 0.5, 0.5)
    A.check()
    tuv, ids, _ = A.trace_rays(s.world, _pick_ray(sc, 0.5, 0.5))
    assert bool(res.hasHit) == (ids[0, 0] != 0xFFFFFFFF)
    if res.hasHit:
        assert res.primID == ids[0, 2]
        K.lib().ospRelease(res.instance)
        K.lib().ospRelease(res.model)


def _pick_ray(sc, x, y):
    cam = sc.camera
    pos = np.array(cam["position"], np.float32)
    d = np.array(cam["direction"], np.float32)
    d = d / np.linalg.norm(d)
    return np.concatenate([pos, d, [0.0, np.inf]])[None, :].astype(np.float32)


@pytest.fixture(scope="module")
def c1_full():
    return scenes.c1()


def test_c1_full_size_first_hit(hip_device, c1_full):
    sc = c1_full
    s = scenes.to_ospray(sc)
    osc = orc.OracleScene(sc)
    rays = _camera_rays(sc, 50000, 5)
    tuv, ids, _ = A.trace_rays(s.world, rays)
    otuv, oids, _ = osc.intersect(rays)
    assert np.array_equal(ids, oids)
    assert np.array_equal(tuv, otuv)
    assert (ids[:, 0] != 0xFFFFFFFF).mean() > 0.5


def test_c1_full_size_frame_properties(hip_device, c1_full):
    sc = c1_full
    s = scenes.to_ospray(sc)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH | K.OSP_FB_ID_INSTANCE | K.OSP_FB_VARIANCE
    fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    for _ in range(2):
        A.render(fb, s.renderer, s.camera, s.world)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    inst = A.map_channel(fb, K.OSP_FB_ID_INSTANCE, sc.width, sc.height)
    assert np.isfinite(col).all()
    assert (col >= 0).all() and (col[..., 3] <= 1).all()
    assert col[..., :3].mean() > 0.01
    assert len(np.unique(inst)) > 50  # terrain plus many sphere instances visible
    assert np.isfinite(K.lib().ospGetVariance(fb.h))
    # a band of rows matches the oracle restricted to those rows
    ofb = orc.OracleFrameBuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    osc = orc.OracleScene(sc)
    cam, ren = orc.make_camera(sc.camera), orc.make_renderer(sc.renderer)
    for _ in range(2):
        ofb.render(osc, cam, ren, 8, (530, 546))
    assert np.array_equal(col[530:546], ofb.channel(K.OSP_FB_COLOR)[530:546])


def test_c_application_runs(hip_device, tmp_path):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "app")
    subprocess.run(["gcc", "-std=c99", "-o", exe, os.path.join(root, "tests", "apps", "pathtracer_app.c"),
                    f"-I{root}/include", f"-L{K.LIBDIR}", "-lospray", f"-Wl,-rpath,{K.LIBDIR}", "-lm"], check=True)
    r = subprocess.run([exe, "--osp:load-modules=hip", "--osp:device=hip", str(tmp_path / "out.ppm"), "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checksum" in r.stdout and int(r.stdout.split()[-1]) > 0
    assert os.path.getsize(tmp_path / "out.ppm") > 160 * 120 * 3


def test_cpp_application_matches_c_application(hip_device, tmp_path):
    """The same scene written against the C++ wrapper renders the same image
    as the C application, byte for byte; its pick hits the red triangles and
    the world's bounds hold the vertices' box."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    link = [f"-I{root}/include", f"-L{K.LIBDIR}", "-lospray", f"-Wl,-rpath,{K.LIBDIR}"]
    c_exe, cpp_exe = str(tmp_path / "c_app"), str(tmp_path / "cpp_app")
    subprocess.run(["gcc", "-std=c99", "-o", c_exe, os.path.join(root, "tests", "apps", "pathtracer_app.c")]
                   + link + ["-lm"], check=True)
    subprocess.run(["g++", "-std=c++17", "-o", cpp_exe, os.path.join(root, "tests", "apps", "pathtracer_cpp_app.cpp")]
                   + link, check=True)
    outs = []
    for exe, name in ((c_exe, "c.ppm"), (cpp_exe, "cpp.ppm")):
        r = subprocess.run([exe, "--osp:load-modules=hip", "--osp:device=hip", str(tmp_path / name), "3"],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs.append(r.stdout.splitlines())
    assert outs[1][0] == outs[0][0]  # frames, variance, checksum
    assert (tmp_path / "c.ppm").read_bytes() == (tmp_path / "cpp.ppm").read_bytes()
    assert outs[1][1] == "pick hit 1 model 0 primID 0"
    # bounds are the BVH's root box: the vertices' box, conservatively padded
    b = np.array([float(x) for x in outs[1][2].split()[1:]])
    box = np.array([-3.0, -1.5, -1.0, 3.0, 1.0, 6.0])
    assert np.all(b[:3] <= box[:3]) and np.all(b[3:] >= box[3:]) and np.allclose(b, box, atol=1e-4)
