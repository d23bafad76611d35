"""Pins the oracle's integrator, BSDFs and lights with closed forms.

  * white furnace: a convex Lambertian object (albedo rho) under an ambient
    light of radiance L reflects exactly rho*L; with NEE and BSDF sampling
    sharing the cosine pdf, each MIS half is deterministic, so every hit
    pixel equals rho*L up to float rounding (pins Lambert eval/sample pdfs,
    ambient sample/eval, the power heuristic and the virtual-light path);
  * a Lambertian plane under a quad light: the estimate converges to
    rho/pi * E with E from Lambert's polygon formula (pins the spherical-
    rectangle sampling, its cosine warp, quad eval and MIS);
  * a plane under a delta sun: exactly rho/pi * E * cos(theta);
  * sRGB colour conversion against the textbook transfer function.
"""
import numpy as np
import pytest

from ospray_amd import scenes
from oracle import orc

POINT_FILTER = 0


def _single(geom, mats, lights, cam, renderer, w=48, h=48):
    return scenes.Scene("t", w, h, [geom], [scenes.Model(0, 0)], [[0]], [scenes.Instance(0)], mats, lights, cam,
                        renderer)


def test_white_furnace():
    unit, faces = scenes.icosphere(2)
    geom = scenes.Geometry("triangles", vertex=unit.astype(np.float32), index=faces.astype(np.uint32))
    sc = _single(geom, [{"type": "obj", "kd": (0.5, 0.5, 0.5)}], [{"type": "ambient", "intensity": 1.0}],
                 {"position": (0.0, 0.0, -4.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 40.0},
                 {"pixelSamples": 1, "maxPathLength": 10})
    fb, _, _ = orc.render_scene(sc, frames=1, fmt=3, channels=1 | 4 | 128)
    col = fb.channel(1)
    hit = fb.channel(128) != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.8
    assert np.abs(col[hit][:, :3] - 0.5).max() < 2e-6
    assert np.abs(col[~hit][:, :3] - 1.0).max() < 2e-6
    assert np.all(col[hit][:, 3] == 1.0) and np.all(col[~hit][:, 3] == 0.0)


def _polygon_irradiance(x, verts, n):
    """Lambert's formula: E/L = 1/2 sum_i angle_i * (n . unit(v_i x v_i+1))."""
    v = verts - x
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    s = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        c = np.cross(a, b)
        s += np.arccos(np.clip(np.dot(a, b), -1, 1)) * np.dot(n, c / np.linalg.norm(c))
    return 0.5 * abs(s)


def test_quad_light_on_plane():
    size = 40.0
    v = np.array([[-size, 0, -size], [size, 0, -size], [size, 0, size], [-size, 0, size]], np.float32)
    geom = scenes.Geometry("quads", vertex=v, index=np.array([[0, 1, 2, 3]], np.uint32))
    L, rho = 10.0, 0.6
    light = {"type": "quad", "position": (-1.0, 3.0, -1.0), "edge1": (2.0, 0.0, 0.0), "edge2": (0.0, 0.0, 2.0),
             "intensity": L}
    cam = {"position": (0.0, 10.0, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 40.0}
    sc = _single(geom, [{"type": "obj", "kd": (rho, rho, rho)}], [light], cam,
                 {"pixelSamples": 16, "maxPathLength": 4, "pixelFilter": POINT_FILTER}, w=16, h=16)
    fb, _, _ = orc.render_scene(sc, frames=8, fmt=3, channels=1 | 4 | 1024)
    col = fb.channel(1)[..., 0]
    pos = fb.channel(1024)
    quad = np.array([[-1, 3, -1], [1, 3, -1], [1, 3, 1], [-1, 3, 1]], np.float64)
    expect = np.array([[rho / np.pi * L * _polygon_irradiance(pos[j, i].astype(np.float64), quad, np.array([0, 1.0, 0]))
                        for i in range(16)] for j in range(16)])
    rel = np.abs(col - expect) / expect
    assert rel.mean() < 0.01
    assert rel.max() < 0.05
    assert abs(col.mean() / expect.mean() - 1) < 0.005


def test_quad_light_emits_on_one_side():
    v = np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32)
    geom = scenes.Geometry("quads", vertex=v, index=np.array([[0, 1, 2, 3]], np.uint32))
    # edge1 x edge2 = +y: emits away from the plane
    light = {"type": "quad", "position": (-1.0, 3.0, -1.0), "edge1": (0.0, 0.0, 2.0), "edge2": (2.0, 0.0, 0.0),
             "intensity": 10.0}
    cam = {"position": (0.0, 10.0, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 30.0}
    sc = _single(geom, [{"type": "obj", "kd": (0.6, 0.6, 0.6)}], [light], cam,
                 {"pixelSamples": 4, "maxPathLength": 4, "pixelFilter": POINT_FILTER}, w=8, h=8)
    fb, _, _ = orc.render_scene(sc, frames=1, fmt=3, channels=1 | 4)
    col = fb.channel(1)[..., :3]
    # the camera above sees the emitting side directly (radiance 10); the
    # plane below the light stays unlit
    seen = np.all(col == 10.0, axis=-1)
    assert 0 < seen.sum() < seen.size
    assert np.abs(col[~seen]).max() == 0.0


def test_delta_sun_on_plane():
    v = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    geom = scenes.Geometry("quads", vertex=v, index=np.array([[0, 1, 2, 3]], np.uint32))
    d = np.array([0.3, -1.0, 0.2])
    d /= np.linalg.norm(d)
    E, rho = 3.0, 0.4
    sun = {"type": "distant", "direction": tuple(d), "intensity": E, "angularDiameter": 0.0}
    cam = {"position": (0.0, 10.0, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 30.0}
    sc = _single(geom, [{"type": "obj", "kd": (rho, rho, rho)}], [sun], cam,
                 {"pixelSamples": 2, "maxPathLength": 4}, w=16, h=16)
    fb, _, _ = orc.render_scene(sc, frames=2, fmt=3, channels=1 | 4)
    expect = rho / np.pi * E * (-np.float32(d[1]))
    assert np.abs(fb.channel(1)[..., :3] - expect).max() < 1e-5


def test_srgb_conversion():
    unit, faces = scenes.icosphere(1)
    geom = scenes.Geometry("triangles", vertex=unit.astype(np.float32), index=faces.astype(np.uint32))
    args = (geom, [{"type": "obj", "kd": (0.7, 0.3, 0.1)}],
            [{"type": "ambient", "intensity": 0.8}, {"type": "distant", "direction": (0, 0, 1.0), "intensity": 1.0}],
            {"position": (0.0, 0.0, -4.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 40.0},
            {"pixelSamples": 1, "maxPathLength": 3})
    lin, _, _ = orc.render_scene(_single(*args), frames=1, fmt=3, channels=1 | 4)
    srgb, _, _ = orc.render_scene(_single(*args), frames=1, fmt=2, channels=1 | 4)
    c = np.clip(lin.channel(1).astype(np.float64), 0, 1)
    ref = np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1 / 2.4) - 0.055)
    ref[..., 3] = c[..., 3]
    assert np.abs(srgb.channel(1).astype(np.float64) - np.floor(ref * 255 + 0.5)).max() <= 1


# ---- emissive geometry (luminous material as a geometry light) -------------------

def _plane_and_emitter(emitter, emit_mat, rho, renderer, cam, w=16, h=16, emitter_xfm=None):
    size = 40.0
    v = np.array([[-size, 0, -size], [size, 0, -size], [size, 0, size], [-size, 0, size]], np.float32)
    plane = scenes.Geometry("quads", vertex=v, index=np.array([[0, 1, 2, 3]], np.uint32))
    insts = [scenes.Instance(0), scenes.Instance(1)]
    if emitter_xfm is not None:
        insts[1] = scenes.Instance(1, xfm=emitter_xfm)
    return scenes.Scene("geo", w, h, [plane, emitter], [scenes.Model(0, 0), scenes.Model(1, 1)], [[0], [1]], insts,
                        [{"type": "obj", "kd": (rho, rho, rho)}, emit_mat], [], cam, renderer)


def test_luminous_quad_mesh_on_plane():
    """Emission of a luminous quad mesh sampled as a geometry light (area
    sampling of both quad halves, area pdf) and hit by BSDF samples, MIS
    combined: the plane converges to rho/pi * L * Lambert's polygon formula.
    With geometry lights off only BSDF-sampled hits count: same limit."""
    L, rho = 10.0, 0.6
    q = np.array([[-1, 3, -1], [1, 3, -1], [1, 3, 1], [-1, 3, 1]], np.float32)
    emitter = scenes.Geometry("quads", vertex=q, index=np.array([[0, 1, 2, 3]], np.uint32))
    cam = {"position": (0.0, 10.0, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 40.0}
    expect = None
    for geo_lights, frames, tol in ((True, 8, 0.01), (False, 32, 0.03)):
        sc = _plane_and_emitter(emitter, {"type": "luminous", "intensity": L}, rho,
                                {"pixelSamples": 16, "maxPathLength": 4, "pixelFilter": POINT_FILTER,
                                 "geometryLights": geo_lights}, cam)
        fb, _, _ = orc.render_scene(sc, frames=frames, fmt=3, channels=1 | 4 | 1024 | 256)
        col = fb.channel(1)[..., 0]
        pos = fb.channel(1024)
        plane = fb.channel(256) == 0
        if expect is None:
            expect = np.array([[rho / np.pi * L * _polygon_irradiance(pos[j, i].astype(np.float64), q.astype(np.float64),
                                                                        np.array([0, 1.0, 0]))
                                for i in range(16)] for j in range(16)])
        assert plane.sum() > 100
        rel = np.abs(col[plane] - expect[plane]) / expect[plane]
        assert abs(col[plane].mean() / expect[plane].mean() - 1) < tol, (geo_lights, col[plane].mean(), expect[plane].mean())
        if geo_lights:
            assert rel.mean() < 0.02
        # the emitter seen directly: radiance L
        assert (fb.channel(256) == 1).any() and np.allclose(col[fb.channel(256) == 1], L)


def test_luminous_sphere_instance_on_plane():
    """A luminous sphere (spheres geometry, instanced with scale 2 and moved
    up): irradiance at plane point x is pi L R^2 h / d^3 for a sphere of
    world radius R whose centre is at height h and distance d."""
    L, rho, r0, s = 4.0, 0.5, 0.5, 2.0
    emitter = scenes.Geometry("spheres", spheres=np.array([[0, 0, 0, r0]], np.float32))
    h = 4.0
    cam = {"position": (0.0, 20.0, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 50.0}
    sc = _plane_and_emitter(emitter, {"type": "luminous", "intensity": L}, rho,
                            {"pixelSamples": 16, "maxPathLength": 3, "pixelFilter": POINT_FILTER}, cam,
                            emitter_xfm=(s, 0, 0, 0, s, 0, 0, 0, s, 0, h, 0))
    fb, _, _ = orc.render_scene(sc, frames=8, fmt=3, channels=1 | 4 | 1024 | 256)
    col = fb.channel(1)[..., 0]
    pos = fb.channel(1024).astype(np.float64)
    plane = fb.channel(256) == 0
    R = r0 * s
    d = np.sqrt(pos[..., 0] ** 2 + h ** 2 + pos[..., 2] ** 2)
    expect = rho / np.pi * np.pi * L * R ** 2 * h / d ** 3
    assert plane.sum() > 100
    assert abs(col[plane].mean() / expect[plane].mean() - 1) < 0.01
    assert (np.abs(col[plane] - expect[plane]) / expect[plane]).mean() < 0.03


def test_semi_transparent_luminous_panel():
    """Camera ray through a luminous panel with transparency t in front of
    the background: exactly L + t * background, alpha 1 - t."""
    L, t, bg = 2.0, 0.25, (0.5, 0.3, 0.1, 0.0)
    q = np.array([[-5, -5, 0], [5, -5, 0], [5, 5, 0], [-5, 5, 0]], np.float32)
    panel = scenes.Geometry("quads", vertex=q, index=np.array([[0, 1, 2, 3]], np.uint32))
    sc = _single(panel, [{"type": "luminous", "intensity": L, "transparency": t}], [],
                 {"position": (0.0, 0.0, -4.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 30.0},
                 {"pixelSamples": 1, "maxPathLength": 5, "backgroundColor": bg}, w=8, h=8)
    fb, _, _ = orc.render_scene(sc, frames=1, fmt=3, channels=1 | 4)
    col = fb.channel(1)
    for k in range(3):
        assert np.all(col[..., k] == np.float32(L + t * bg[k]))
    assert np.all(col[..., 3] == np.float32(1.0 - t))
