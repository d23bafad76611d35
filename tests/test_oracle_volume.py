"""Pins of the oracle's volume path (structured-regular grid, transfer
function, macrocell majorants, delta tracking, Henyey-Greenstein phase) by
closed forms, independent library routines and invariants.

The reference samples volumes through OpenVKL 2.0.1 (not in the reference
tree); the grid and interval definitions the oracle fixes for it are pinned
here by scipy's trilinear interpolation, slab geometry and the exponential
law of homogeneous and linear media.
"""
import math

import numpy as np
import pytest

from ospray_amd import scenes as S
from oracle import orc

NEXT_BEFORE_1 = np.float32(0.99999994)


def _vol_scene(vol, xfm=(1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0), inst_id=None, width=32, height=32, camera=None,
               lights=(), renderer=None, extra=None):
    """A world of one instance holding only `vol`; `extra` = (geometries,
    models, materials) added as a second instance."""
    geoms, models, mats, groups, insts = [], [], [], [[]], [S.Instance(0, xfm=xfm, user_id=inst_id)]
    gvols = [[0]]
    if extra is not None:
        g2, m2, mat2 = extra
        geoms += g2
        models += m2
        mats += mat2
        groups.append(list(range(len(m2))))
        gvols.append([])
        insts.append(S.Instance(1))
    cam = camera or {"position": (0.0, 0.0, -10.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0),
                     "fovy": 30.0, "aspect": width / height}
    return S.Scene("vol", width, height, geoms, models, groups, insts, mats, list(lights), cam,
                   dict(renderer or {"pixelSamples": 1, "maxPathLength": 5}), volumes=[vol], group_volumes=gvols)


def _const_volume(value=0.5, opacity=0.4, dims=(6, 5, 4), origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0),
                  density=1.0, color=(0.0, 0.0, 0.0), **kw):
    nx, ny, nz = dims
    return S.Volume(np.full((nz, ny, nx), value, np.float32), tf_color=np.array([color, color], np.float32),
                    tf_opacity=np.array([opacity, opacity], np.float32), origin=origin, spacing=spacing,
                    density_scale=density, **kw)


# ---- grid sampling and transfer function ---------------------------------------------------

def test_trilinear_matches_scipy():
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(5)
    data = rng.random((5, 7, 9), dtype=np.float32)  # (nz, ny, nx)
    origin, spacing = (1.0, -2.0, 0.5), (0.5, 0.25, 2.0)
    vol = S.Volume(data, np.zeros((2, 3), np.float32), np.zeros(2, np.float32), origin=origin, spacing=spacing)
    os_ = orc.OracleScene(_vol_scene(vol))
    idx = rng.random((4000, 3)) * (np.array([8, 6, 4]) - 1e-3)  # x, y, z index space, inside
    pts = (np.array(origin) + idx * np.array(spacing)).astype(np.float32)
    got = os_.volume_sample(0, pts)
    lidx = (pts.astype(np.float64) - np.array(origin)) / np.array(spacing)
    ref = map_coordinates(data.astype(np.float64), [lidx[:, 2], lidx[:, 1], lidx[:, 0]], order=1)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    # grid corners return the voxel itself; points outside the box are NaN
    corners = np.array([[1.0, -2.0, 0.5], [1.0 + 8 * 0.5, -2.0 + 6 * 0.25, 0.5 + 4 * 2.0]], np.float32)
    np.testing.assert_array_equal(os_.volume_sample(0, corners), [data[0, 0, 0], data[4, 6, 8]])
    outside = np.array([[0.9, 0.0, 1.0], [2.0, -2.1, 1.0], [2.0, -1.0, 8.6]], np.float32)
    assert np.isnan(os_.volume_sample(0, outside)).all()


def test_transfer_function_matches_interp():
    rng = np.random.default_rng(6)
    col = rng.random((5, 3), dtype=np.float32)
    op = rng.random(7, dtype=np.float32)
    vol = S.Volume(np.zeros((2, 2, 2), np.float32), col, op, tf_range=(-1.0, 3.0))
    os_ = orc.OracleScene(_vol_scene(vol))
    v = np.concatenate([np.linspace(-2.0, 4.0, 3001), [-1.0, 3.0, 1.0]]).astype(np.float32)
    got = os_.tf_get(0, v)
    u = np.clip((v.astype(np.float64) + 1.0) / 4.0, 0.0, float(NEXT_BEFORE_1))
    for c in range(3):
        np.testing.assert_allclose(got[:, c], np.interp(u, np.linspace(0, 1, 5), col[:, c]), rtol=1e-5, atol=2e-6)
    np.testing.assert_allclose(got[:, 3], np.interp(u, np.linspace(0, 1, 7), op), rtol=1e-5, atol=2e-6)


def test_max_opacity_is_a_majorant():
    rng = np.random.default_rng(7)
    op = rng.random(9, dtype=np.float32)
    vol = S.Volume(np.zeros((2, 2, 2), np.float32), np.zeros((2, 3), np.float32), op, tf_range=(2.0, 6.0))
    os_ = orc.OracleScene(_vol_scene(vol))
    a = rng.uniform(1.0, 7.0, 500)
    b = a + rng.uniform(0.0, 3.0, 500)
    mo = os_.tf_max_opacity(0, np.stack([a, b], 1))
    for i in range(500):
        vals = np.linspace(a[i], b[i], 400).astype(np.float32)
        assert os_.tf_get(0, vals)[:, 3].max() <= mo[i] + 1e-6
    # the whole range gives the largest control point, a point range the local maximum
    assert os_.tf_max_opacity(0, [[2.0, 6.0]])[0] == op.max()
    assert os_.tf_max_opacity(0, [[2.0, 2.0]])[0] == op[0]


def test_macrocell_intervals_cover_the_box_and_bound_the_samples():
    rng = np.random.default_rng(8)
    dims = (40, 33, 20)
    data = rng.random(dims[::-1], dtype=np.float32)
    origin, spacing = np.array([-1.0, 0.5, 2.0]), np.array([0.1, 0.07, 0.2])
    vol = S.Volume(data, np.zeros((2, 3), np.float32), np.ones(2, np.float32), origin=tuple(origin),
                   spacing=tuple(spacing))
    os_ = orc.OracleScene(_vol_scene(vol))
    upper = origin + (np.array(dims) - 1) * spacing
    centre = 0.5 * (origin + upper)
    checked = 0
    for _ in range(200):
        o = centre + rng.normal(size=3) * 6.0
        target = origin + rng.random(3) * (upper - origin)
        d = target - o
        d /= np.linalg.norm(d)
        iv = os_.volume_intervals(0, np.concatenate([o, d, [0.0, np.inf]]))
        inv = 1.0 / d
        t_a, t_b = (origin - o) * inv, (upper - o) * inv
        entry = max(np.minimum(t_a, t_b).max(), 0.0)
        exit_ = np.maximum(t_a, t_b).min()
        assert len(iv) > 0
        assert abs(iv[0, 0] - entry) <= 1e-4 * max(1.0, entry)
        assert abs(iv[-1, 1] - exit_) <= 1e-4 * max(1.0, exit_)
        np.testing.assert_array_equal(iv[1:, 0], iv[:-1, 1])  # contiguous
        assert (iv[:, 1] > iv[:, 0]).all()
        u = rng.uniform(0.05, 0.95, (len(iv), 24))
        t = iv[:, :1] + u * (iv[:, 1:2] - iv[:, :1])
        pts = (o[None, None, :] + t[:, :, None] * d[None, None, :]).reshape(-1, 3).astype(np.float32)
        smp = os_.volume_sample(0, pts).reshape(len(iv), 24)
        ok = ~np.isnan(smp)
        lo = np.broadcast_to(iv[:, 2:3], smp.shape)
        hi = np.broadcast_to(iv[:, 3:4], smp.shape)
        assert (smp[ok] >= lo[ok] - 1e-6).all() and (smp[ok] <= hi[ok] + 1e-6).all()
        checked += ok.sum()
    assert checked > 5000


# ---- delta tracking: closed forms ----------------------------------------------------------

def _binom_tol(p, n, k=5.0):
    return k * math.sqrt(max(p * (1 - p), 1e-4) / n)


def test_homogeneous_transmittance_and_free_path():
    mu = 0.8
    vol = _const_volume(opacity=mu, dims=(6, 5, 4), origin=(0.0, 0.0, 0.0))  # box [0,5]x[0,4]x[0,3]
    os_ = orc.OracleScene(_vol_scene(vol))
    diag = np.array([5.0, 4.0, 3.0]) / np.linalg.norm([5.0, 4.0, 3.0])
    rays = np.array([[-1, 2, 1.5, 1, 0, 0, 0, np.inf],  # along x: length 5
                     [2.5, -3, 1.0, 0, 1, 0, 0, np.inf],  # along y: length 4
                     [2.5, 2.0, -1, 0, 0, 1, 0, 2.5],  # along z, stopped after 1.5 inside
                     [*-diag, *diag, 0, np.inf]], np.float32)  # corner to corner
    lengths = np.array([5.0, 4.0, 1.5, float(np.linalg.norm([5, 4, 3]))])
    t_entry = np.array([1.0, 3.0, 1.0, 1.0])
    n = 20000
    T, hit, md = os_.volume_estimates(rays, n)
    for i, L in enumerate(lengths):
        p = math.exp(-mu * L)
        assert abs(T[i] - p) <= _binom_tol(p, n), (i, T[i], p)
        assert abs((1 - hit[i]) - p) <= _binom_tol(p, n), (i, hit[i], p)
        # mean distance to the collision, given one: entry + E[d | d < L]
        expect = t_entry[i] + 1.0 / mu - L * p / (1.0 - p)
        assert abs(md[i] - expect) < 0.03 * expect, (i, md[i], expect)


def test_linear_medium_null_collisions_and_instance_scale():
    # value = x / 4 over x in [0, 4]; opacity = value: mu(x) = 2 * x / 4
    nx = 9
    x = np.linspace(0.0, 1.0, nx, dtype=np.float32)
    data = np.broadcast_to(x[None, None, :], (3, 3, nx)).copy()
    vol = S.Volume(data, np.zeros((2, 3), np.float32), np.array([0.0, 1.0], np.float32), spacing=(0.5, 1.0, 1.0),
                   density_scale=2.0)
    n = 20000
    for scale in (1.0, 2.0):
        xfm = (scale, 0, 0, 0, scale, 0, 0, 0, scale, 0.0, 0.0, 0.0)
        os_ = orc.OracleScene(_vol_scene(vol, xfm=xfm))
        s = scale
        rays = np.array([[-1, 1 * s, 1 * s, 1, 0, 0, 0, np.inf],  # along x through the gradient
                         [1 * s, -1, 1 * s, 0, 1, 0, 0, np.inf],  # along y at x = 1: mu = 0.5
                         [3 * s, 1 * s, -1, 0, 0, 1, 0, np.inf]], np.float32)  # along z at x = 3: mu = 1.5
        # optical depth per unit of the world ray parameter: the instance's
        # scale stretches every path through the grid
        tau = np.array([2.0 * 4.0 / 2.0, 0.5 * 2.0, 1.5 * 2.0]) * s
        T, hit, _ = os_.volume_estimates(rays, n, seed=11)
        for i in range(3):
            p = math.exp(-tau[i])
            assert abs(T[i] - p) <= _binom_tol(p, n), (scale, i, T[i], p)
            assert abs((1 - hit[i]) - p) <= _binom_tol(p, n), (scale, i, hit[i], p)


def test_transmittance_across_many_macrocells():
    """Tracks through 6 x 2 x 2 macrocells (16^3 cells each): the exponential
    draw of the distance to the next tentative collision carries over from
    one macrocell interval to the next, so a homogeneous medium still passes
    exp(-mu L) and a ramp whose majorant changes from cell to cell passes
    exp(-integral of mu)."""
    n = 20000
    nx, ny, nz = 97, 33, 33  # spacing 0.125: box [0,12] x [0,4] x [0,4]
    mu = 0.3
    hom = _const_volume(opacity=mu, dims=(nx, ny, nz), spacing=(0.125, 0.125, 0.125))
    d = np.array([12.0, 4.0, 0.0]) / math.hypot(12.0, 4.0)
    rays = np.array([[-1, 2.1, 1.9, 1, 0, 0, 0, np.inf],  # along x: 6 macrocells, length 12
                     [*(np.array([0.0, 0.0, 1.3]) - d), *d, 0, np.inf]], np.float32)  # x-y diagonal
    lengths = [12.0, math.hypot(12.0, 4.0)]
    T, hit, _ = orc.OracleScene(_vol_scene(hom)).volume_estimates(rays, n, seed=3)
    for i, L in enumerate(lengths):
        p = math.exp(-mu * L)
        assert abs(T[i] - p) <= _binom_tol(p, n), (i, T[i], p)
        assert abs((1 - hit[i]) - p) <= _binom_tol(p, n), (i, hit[i], p)
    # value x / 12, opacity = value, densityScale 0.5: mu(x) = x / 24
    x = np.linspace(0.0, 1.0, nx, dtype=np.float32)
    data = np.broadcast_to(x[None, None, :], (nz, ny, nx)).copy()
    ramp = S.Volume(data, np.zeros((2, 3), np.float32), np.array([0.0, 1.0], np.float32),
                    spacing=(0.125, 0.125, 0.125), density_scale=0.5)
    rays = np.array([[-1, 2.1, 1.9, 1, 0, 0, 0, np.inf],  # the whole ramp: tau = 3
                     [13, 0.7, 3.3, -1, 0, 0, 0, np.inf],  # the other way
                     [-1, 1.0, 1.0, 1, 0, 0, 0, 7.0]], np.float32)  # x in [0, 6]: tau = 0.75
    tau = [3.0, 3.0, 0.75]
    T, hit, _ = orc.OracleScene(_vol_scene(ramp)).volume_estimates(rays, n, seed=5)
    for i in range(3):
        p = math.exp(-tau[i])
        assert abs(T[i] - p) <= _binom_tol(p, n), (i, T[i], p)


# ---- Henyey-Greenstein -------------------------------------------------------------------------

def _grid_samples(k=256):
    u = (np.arange(k) + 0.5) / k
    a, b = np.meshgrid(u, u)
    return np.stack([a.ravel(), b.ravel()], 1).astype(np.float32)


@pytest.mark.parametrize("g", [0.0, 0.5, -0.7, 0.95])
def test_hg_sampling_matches_its_density(g):
    wo = np.array([0.3, -0.5, 0.81], np.float32)
    wo /= np.linalg.norm(wo)
    s2 = _grid_samples()
    wi, pdf, typ = orc.hg_sample(g, wo, s2)
    np.testing.assert_allclose(np.linalg.norm(wi, axis=1), 1.0, atol=2e-6)
    assert (typ == (1 << 1 | 1 << 2 | 1 << 4 | 1 << 5)).all()  # SCATTERING_SMOOTH
    ev = orc.hg_eval(g, wo, wi)
    np.testing.assert_allclose(ev, pdf, rtol=2e-3 if abs(g) < 0.9 else 2e-2)
    cos = wi.astype(np.float64) @ -wo.astype(np.float64)
    assert abs(cos.mean() - g) < 2e-3  # E[cos theta] = g
    # the phase function integrates to one over the sphere (Fibonacci points)
    m = 200000
    k = np.arange(m) + 0.5
    z = 1 - 2 * k / m
    ph = math.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1).astype(np.float32)
    assert abs(orc.hg_eval(g, wo, dirs).astype(np.float64).mean() * 4 * math.pi - 1.0) < 2e-3
    if g == 0.0:
        np.testing.assert_allclose(pdf, 1.0 / (4.0 * math.pi), rtol=1e-6)


def test_hg_delta_limit():
    wo = np.array([0.0, 0.0, 1.0], np.float32)
    wi, pdf, typ = orc.hg_sample(0.9995, wo, [[0.3, 0.7]])
    np.testing.assert_array_equal(wi[0], -wo)
    assert np.isinf(pdf[0]) and typ[0] == 1 << 3  # specular transmission
    wi, pdf, typ = orc.hg_sample(-0.9995, wo, [[0.3, 0.7]])
    np.testing.assert_array_equal(wi[0], wo)
    assert np.isinf(pdf[0]) and typ[0] == 1 << 0  # specular reflection


# ---- the path tracer with volumes ------------------------------------------------------------

def _c0_small():
    return S.c0(width=48, height=48, spp=1, max_path_length=4)


def test_zero_opacity_volume_changes_nothing():
    """A volume whose transfer function is transparent everywhere draws no
    random numbers and must leave the image bit-identical."""
    base = _c0_small()
    fb0, _, _ = orc.render_scene(base, frames=2)
    sc = _c0_small()
    sc.volumes = [_const_volume(opacity=0.0, dims=(8, 8, 8), origin=(-4.0, -4.0, -4.0), spacing=(1.0, 1.0, 1.0))]
    sc.groups.append([])
    sc.group_volumes = [[] for _ in sc.groups[:-1]] + [[0]]
    sc.instances.append(S.Instance(len(sc.groups) - 1))
    fb1, _, _ = orc.render_scene(sc, frames=2)
    np.testing.assert_array_equal(fb0.channel(1), fb1.channel(1))


def test_absorbing_box_in_front_of_background():
    """Camera rays through a black homogeneous box towards a white
    background: each sample passes with probability exp(-mu d)."""
    mu, d = 0.6, 2.0
    vol = _const_volume(opacity=mu, dims=(21, 21, 3), origin=(-10.0, -10.0, 0.0), spacing=(1.0, 1.0, 1.0))
    cam = {"position": (0.0, 0.0, -10.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 2.0,
           "aspect": 1.0}
    sc = _vol_scene(vol, width=32, height=32, camera=cam,
                    renderer={"pixelSamples": 16, "maxPathLength": 3, "backgroundColor": (1.0, 1.0, 1.0, 1.0)})
    fb, _, _ = orc.render_scene(sc, frames=1)
    img = fb.channel(1)[..., 0].astype(np.float64)
    p = math.exp(-mu * d)
    n = img.size * 16
    assert abs(img.mean() - p) <= _binom_tol(p, n) + 2e-4, (img.mean(), p)
    # camera rays that collide report the volume in the ID buffers
    fb2, _, _ = orc.render_scene(sc, frames=1, channels=1 | 2 | 64 | 128 | 256)
    assert set(np.unique(fb2.channel(128))) <= {0, 0xFFFFFFFF}  # model id: group-local index 0
    assert set(np.unique(fb2.channel(256))) <= {0, 0xFFFFFFFF}


def test_shadow_rays_through_an_absorbing_slab():
    """A Lambertian plane lit by a distant light straight above; an
    absorbing slab above the camera covers half the plane.  Under the slab
    the direct light is attenuated by exp(-mu d) on average."""
    mu, d = 0.7, 1.0
    kd, E = 0.8, 2.0
    plane = S.Geometry("quads", vertex=np.array([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]], np.float32),
                       index=np.array([[0, 3, 2, 1]], np.uint32))
    vol = _const_volume(opacity=mu, dims=(11, 2, 6), origin=(-5.0, 1.0, 0.0), spacing=(1.0, d, 1.0))
    cam = {"position": (0.0, 0.8, 0.0), "direction": (0.0, -1.0, 0.0), "up": (0.0, 0.0, 1.0), "fovy": 60.0,
           "aspect": 1.0}
    sc = _vol_scene(vol, width=64, height=64, camera=cam,
                    lights=[{"type": "distant", "direction": (0.0, -1.0, 0.0), "intensity": E}],
                    renderer={"pixelSamples": 8, "maxPathLength": 1},
                    extra=([plane], [S.Model(0, 0)], [{"type": "obj", "kd": (kd, kd, kd)}]))
    fb, _, _ = orc.render_scene(sc, frames=1)
    img = fb.channel(1)[..., 0].astype(np.float64)
    L0 = kd / math.pi * E
    # image rows run along the camera's up (+z) axis; keep away from z = 0
    lit, shaded = img[:24], img[40:]
    if lit.mean() < shaded.mean():
        lit, shaded = shaded, lit
    np.testing.assert_allclose(lit, L0, rtol=1e-5)
    p = math.exp(-mu * d)
    n = shaded.size * 8
    assert abs(shaded.mean() / L0 - p) <= _binom_tol(p, n), (shaded.mean() / L0, p)
