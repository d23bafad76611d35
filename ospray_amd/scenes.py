"""Seeded synthetic scenes of the benchmark configurations.

All randomness comes from PCG32 (O'Neill's XSH-RR generator, seeded with
pcg32_srandom(seed, 54)); uniform floats are (u32 >> 8) * 2^-24.

  c0(): 512x512, 1 spp, maxPathLength 5; 64 icospheres (subdivision 3) with
        radius U[0.2,0.6], centre U[-5,5]^3 and Lambert kd U[0.1,0.9]^3 plus
        a 20x20 ground quad at y=-6; one quad light (intensity 10) and an
        ambient light (0.2).  Seed 42.
  c4(): 1280x720, 2 spp, maxPathLength 10; the C0 spheres inside a 512^3
        fBm density volume (seed 45) with an 8-point transfer function.
  c1(): 1920x1080, 4 spp, maxPathLength 8; 1448^2 fBm heightfield over
        [-50,50]^2 (4,187,618 triangles, 5 octaves, amplitude 2) plus 256
        instances of one subdivision-4 icosphere; 60/25/15 % diffuse / metal
        (roughness U[0,0.5]) / glass (eta 1.5) per instance; a sun
        (distant, 0.53 deg) and 4 quad lights.  Seed 43.

A Scene is a plain description (numpy arrays and dicts); `to_ospray` builds
it through the C API and returns the committed handles.
"""
from __future__ import annotations

import dataclasses
import os
import math
from typing import Optional

import numpy as np

_MASK64 = (1 << 64) - 1


class Pcg32:
    MULT = 6364136223846793005

    def __init__(self, seed: int, seq: int = 54):
        self.state = 0
        self.inc = ((seq << 1) | 1) & _MASK64
        self.next_u32()
        self.state = (self.state + seed) & _MASK64
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * self.MULT + self.inc) & _MASK64
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u32() >> 8) * (1.0 / 16777216.0)
        return lo + (hi - lo) * u


def pcg32_draws(rng: Pcg32, n: int) -> np.ndarray:
    """The next n outputs of `rng` as a uint32 array (the same values as n
    calls of rng.next_u32(), which advances rng accordingly): the LCG states
    s_k = a^k s_0 + c (a^(k-1) + ... + 1) are formed by doubling, so the
    whole run costs O(log n) vector steps."""
    a, c = np.uint64(Pcg32.MULT), np.uint64(rng.inc)
    A = np.ones(1, np.uint64)  # a^k for k in [0, m)
    Cs = np.zeros(1, np.uint64)  # c (a^(k-1) + ... + 1)
    am, cm = a, c  # a^m, c (a^(m-1) + ... + 1) for the current block size m
    with np.errstate(over="ignore"):
        while A.size < n:
            A = np.concatenate([A, A * am])
            Cs = np.concatenate([Cs, Cs * am + cm])
            cm = cm * am + cm
            am = am * am
        A, Cs = A[:n], Cs[:n]
        old = A * np.uint64(rng.state) + Cs
        xorshifted = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = old >> np.uint64(59)
        out = ((xorshifted >> rot) | (xorshifted << ((np.uint64(32) - rot) & np.uint64(31)))) & np.uint64(0xFFFFFFFF)
        last = old[-1] if n else np.uint64(rng.state)
        if n:
            rng.state = int(last * a + c) & _MASK64
    return out.astype(np.uint32)


def _fade_np(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def value_fbm3(n: int, rng: Pcg32, octaves: int = 6, base_period: int = 128) -> np.ndarray:
    """n^3 fBm of seeded lattice value noise, float32 (nz, ny, nx).

    Octave o has lattice period base_period / 2^o voxels, gain 0.5^o and
    lattice values 2u - 1 (u from the PCG32 draws, z-major then y then x);
    between lattice points the noise is interpolated per axis with the
    quintic fade 6t^5 - 15t^4 + 10t^3, separably (x, then y, then z), in
    float32.  Slabs of z are summed in parallel threads."""
    import concurrent.futures as cf

    i = np.arange(n)
    planes = []  # per octave: (lattice x/y-interpolated planes, period, amplitude)
    amp = 1.0
    for o in range(octaves):
        per = max(1, base_period >> o)
        L = (n - 1) // per + 2
        u = pcg32_draws(rng, L * L * L).astype(np.float64) * (1.0 / 4294967296.0)
        lat = (2.0 * u - 1.0).astype(np.float32).reshape(L, L, L)
        c = i // per
        w = _fade_np((i % per) / per).astype(np.float32)
        v = lat[:, :, c] * (1.0 - w) + lat[:, :, c + 1] * w  # x
        vy = np.empty((L, n, n), np.float32)
        for k in range(L):  # y
            vy[k] = v[k, c, :] * (1.0 - w)[:, None] + v[k, c + 1, :] * w[:, None]
        planes.append((vy, per, np.float32(amp)))
        amp *= 0.5
    total = np.empty((n, n, n), np.float32)

    def slab(z0, z1):
        acc = np.zeros((z1 - z0, n, n), np.float32)
        for vy, per, a in planes:
            for z in range(z0, z1):
                k, f = divmod(z, per)
                wz = np.float32(_fade_np(f / per))
                acc[z - z0] += a * (vy[k] * (np.float32(1.0) - wz) + vy[k + 1] * wz)
        total[z0:z1] = acc

    step = 16
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda z0: slab(z0, min(n, z0 + step)), range(0, n, step)))
    return total


@dataclasses.dataclass
class Geometry:
    type: str  # "triangles" | "quads" | "spheres"
    vertex: Optional[np.ndarray] = None  # (N,3) float32
    index: Optional[np.ndarray] = None  # (M,3|4) uint32
    normal: Optional[np.ndarray] = None  # (N,3) float32
    spheres: Optional[np.ndarray] = None  # (N,4) float32 centre+radius

    @property
    def num_prims(self) -> int:
        return int(self.spheres.shape[0] if self.type == "spheres" else self.index.shape[0])


@dataclasses.dataclass
class Model:
    geometry: int
    material: int = -1
    user_id: Optional[int] = None


@dataclasses.dataclass
class Instance:
    group: int
    xfm: tuple = (1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)
    user_id: Optional[int] = None


@dataclasses.dataclass
class Volume:
    """A structured-regular volume (vertex-centred voxels) with its
    piecewise-linear transfer function and volumetric-model parameters."""
    data: np.ndarray  # (nz, ny, nx) float32, x fastest
    tf_color: np.ndarray  # (N, 3) float32
    tf_opacity: np.ndarray  # (M,) float32
    origin: tuple = (0.0, 0.0, 0.0)
    spacing: tuple = (1.0, 1.0, 1.0)
    tf_range: tuple = (0.0, 1.0)
    density_scale: float = 1.0
    anisotropy: float = 0.0
    user_id: Optional[int] = None


@dataclasses.dataclass
class Scene:
    name: str
    width: int
    height: int
    geometries: list
    models: list
    groups: list  # list of lists of model indices
    instances: list
    materials: list  # dicts with "type" + OSPRay parameters
    lights: list  # dicts with "type" + OSPRay parameters
    camera: dict
    renderer: dict
    volumes: list = dataclasses.field(default_factory=list)  # Volume
    group_volumes: list = dataclasses.field(default_factory=list)  # per group: volume indices

    @property
    def num_triangles(self) -> int:
        n = 0
        for inst in self.instances:
            for m in self.groups[inst.group]:
                g = self.geometries[self.models[m].geometry]
                n += g.num_prims * (2 if g.type == "quads" else 1) if g.type != "spheres" else 0
        return n


# ---- geometry generators -----------------------------------------------------------------

def icosphere(subdiv: int):
    """Unit icosphere: (vertices (N,3) float64, triangles (M,3) int)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(v), np.array(faces, np.int64)


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


class GradientNoise2D:
    """Seeded 2D gradient (Perlin-style) noise, vectorised."""

    def __init__(self, rng: Pcg32):
        perm = list(range(256))
        for i in range(255, 0, -1):  # Fisher-Yates with PCG draws
            j = rng.next_u32() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        self.perm = np.array(perm + perm, np.int64)
        ang = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(256)])
        self.grad = np.stack([np.cos(ang), np.sin(ang)], axis=1)

    def __call__(self, x, y):
        xi = np.floor(x).astype(np.int64)
        yi = np.floor(y).astype(np.int64)
        xf, yf = x - xi, y - yi
        xi &= 255
        yi &= 255
        p = self.perm

        def g(ix, iy, dx, dy):
            h = p[p[ix] + iy] & 255
            gr = self.grad[h]
            return gr[..., 0] * dx + gr[..., 1] * dy

        n00 = g(xi, yi, xf, yf)
        n10 = g(xi + 1, yi, xf - 1, yf)
        n01 = g(xi, yi + 1, xf, yf - 1)
        n11 = g(xi + 1, yi + 1, xf - 1, yf - 1)
        u, v = _fade(xf), _fade(yf)
        a = n00 + u * (n10 - n00)
        b = n01 + u * (n11 - n01)
        return a + v * (b - a)


def fbm(noise, x, y, octaves=5, lacunarity=2.0, gain=0.5, freq=1.0 / 20.0):
    s = np.zeros_like(x)
    amp = 1.0
    for _ in range(octaves):
        s += amp * noise(x * freq, y * freq)
        freq *= lacunarity
        amp *= gain
    return s


def heightfield(n: int, extent: float, amplitude: float, noise):
    """n x n vertex grid over [-extent, extent]^2, y = amplitude * fBm."""
    c = np.linspace(-extent, extent, n)
    X, Z = np.meshgrid(c, c)  # Z varies along rows (j), X along columns (i)
    Y = amplitude * fbm(noise, X, Z)
    vert = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i = np.arange(n - 1)
    j = np.arange(n - 1)
    I, J = np.meshgrid(i, j)
    v00 = (J * n + I).ravel()
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    # (v00, v01, v11) and (v00, v11, v10): geometric normals point to +y
    tri = np.empty((2 * v00.size, 3), np.uint32)
    tri[0::2] = np.stack([v00, v01, v11], axis=1)
    tri[1::2] = np.stack([v00, v11, v10], axis=1)
    return vert, tri


# ---- configurations --------------------------------------------------------------------

def _look_at(pos, target=(0.0, 0.0, 0.0)):
    return tuple(float(t - p) for p, t in zip(pos, target))


def c0(width: int = 512, height: int = 512, spp: int = 1, max_path_length: int = 5, seed: int = 42) -> Scene:
    rng = Pcg32(seed)
    unit, faces = icosphere(3)
    idx = faces.astype(np.uint32)
    geoms, models, mats = [], [], []
    for i in range(64):
        r = rng.uniform(0.2, 0.6)
        c = np.array([rng.uniform(-5, 5), rng.uniform(-5, 5), rng.uniform(-5, 5)])
        kd = (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9))
        geoms.append(Geometry("triangles", vertex=(c + r * unit).astype(np.float32), index=idx,
                              normal=unit.astype(np.float32)))
        mats.append({"type": "obj", "kd": kd})
        models.append(Model(geometry=i, material=i))
    ground = np.array([[-10, -6, -10], [10, -6, -10], [10, -6, 10], [-10, -6, 10]], np.float32)
    geoms.append(Geometry("quads", vertex=ground, index=np.array([[0, 1, 2, 3]], np.uint32)))
    mats.append({"type": "obj"})
    models.append(Model(geometry=64, material=64))
    lights = [
        # emission side is edge1 x edge2 (QuadLight.ispc:23-25): -y, onto the scene
        {"type": "quad", "position": (-2.0, 8.0, -2.0), "edge1": (4.0, 0.0, 0.0), "edge2": (0.0, 0.0, 4.0),
         "intensity": 10.0},
        {"type": "ambient", "intensity": 0.2},
    ]
    pos = (0.0, 4.0, -20.0)
    return Scene("c0", width, height, geoms, models, [list(range(65))], [Instance(0)], mats, lights,
                 camera={"position": pos, "direction": _look_at(pos), "up": (0.0, 1.0, 0.0), "fovy": 45.0,
                         "aspect": width / height},
                 renderer={"pixelSamples": spp, "maxPathLength": max_path_length})


def c1(width: int = 1920, height: int = 1080, spp: int = 4, max_path_length: int = 8, seed: int = 43,
       grid: int = 1448, instances: int = 256) -> Scene:
    rng = Pcg32(seed)
    noise = GradientNoise2D(rng)
    vert, tri = heightfield(grid, 50.0, 2.0, noise)
    unit, faces = icosphere(4)
    geoms = [Geometry("triangles", vertex=vert, index=tri),
             Geometry("triangles", vertex=unit.astype(np.float32), index=faces.astype(np.uint32),
                      normal=unit.astype(np.float32))]
    mats = [{"type": "obj", "kd": (0.55, 0.5, 0.45)}]
    models = [Model(geometry=0, material=0)]
    groups = [[0]]
    insts = [Instance(0)]
    for k in range(instances):
        u = rng.uniform()
        if u < 0.60:
            mats.append({"type": "obj", "kd": (rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9))})
        elif u < 0.85:
            mats.append({"type": "metal", "roughness": rng.uniform(0.0, 0.5)})
        else:
            mats.append({"type": "glass", "eta": 1.5})
        s = rng.uniform(0.5, 2.5)
        x, z = rng.uniform(-45, 45), rng.uniform(-45, 45)
        y = float(2.0 * fbm(noise, np.array([x]), np.array([z]))[0]) + s + rng.uniform(0.0, 3.0)
        models.append(Model(geometry=1, material=len(mats) - 1))
        groups.append([len(models) - 1])
        insts.append(Instance(len(groups) - 1, xfm=(s, 0, 0, 0, s, 0, 0, 0, s, x, y, z)))
    d = np.array([-0.3, -1.0, 0.4])
    lights = [{"type": "distant", "direction": tuple(float(v) for v in d / np.linalg.norm(d)), "intensity": 2.5,
               "angularDiameter": 0.53, "color": (1.0, 0.95, 0.85)}]
    for k, (qx, qz) in enumerate([(-20, -20), (20, -20), (-20, 20), (20, 20)]):
        lights.append({"type": "quad", "position": (qx - 3.0, 25.0, qz - 3.0), "edge1": (6.0, 0.0, 0.0),
                       "edge2": (0.0, 0.0, 6.0), "intensity": 4.0,
                       "color": (1.0, 0.9 + 0.03 * k, 0.8 + 0.05 * k)})
    pos = (0.0, 20.0, -60.0)
    return Scene("c1", width, height, geoms, models, groups, insts, mats, lights,
                 camera={"position": pos, "direction": _look_at(pos), "up": (0.0, 1.0, 0.0), "fovy": 40.0,
                         "aspect": width / height},
                 renderer={"pixelSamples": spp, "maxPathLength": max_path_length})


def c2_views(n: int = 64, seed: int = 44, radius: float = 30.0) -> list:
    """The C2 batch: n views of the C1 scene with positions uniform on a
    sphere of `radius` about the origin (z = 2u - 1, phi = 2 pi v), looking
    at the origin, resolution drawn from {256^2, 512^2, 768x432, 1024^2} and
    spp from {1, 2, 4, 8}; maxPathLength 8, fovy 40.  Views below the
    terrain still render (they see its underside)."""
    rng = Pcg32(seed)
    sizes = [(256, 256), (512, 512), (768, 432), (1024, 1024)]
    views = []
    for _ in range(n):
        z = 2.0 * rng.uniform() - 1.0
        phi = 2.0 * math.pi * rng.uniform()
        r = math.sqrt(max(0.0, 1.0 - z * z))
        pos = (radius * r * math.cos(phi), radius * z, radius * r * math.sin(phi))
        w, h = sizes[rng.next_u32() % 4]
        spp = (1, 2, 4, 8)[rng.next_u32() % 4]
        up = (0.0, 1.0, 0.0) if abs(z) < 0.999 else (1.0, 0.0, 0.0)
        views.append({"width": w, "height": h, "spp": spp,
                      "camera": {"position": pos, "direction": _look_at(pos), "up": up, "fovy": 40.0,
                                 "aspect": w / h}})
    return views


def c4(width: int = 1280, height: int = 720, spp: int = 2, max_path_length: int = 10, seed: int = 45,
       grid: int = 512, base_period: int = 128) -> Scene:
    """The C0 spheres, ground and lights embedded in a grid^3 fBm density
    volume over [-6, 6]^3: value-noise fBm (6 octaves, lattice period
    base_period voxels halving per octave, seed 45) mapped to
    clamp(0.5 + fBm, 0, 1); piecewise-linear transfer function of 8 points
    over [0, 1] with seeded colours U[0.3, 1]^3 and opacity 0.5 * k / 7
    (opacity scale 0.5), densityScale 1, isotropic scattering."""
    sc = c0(width, height, spp=spp, max_path_length=max_path_length)
    rng = Pcg32(seed)
    dens = np.clip(0.5 + value_fbm3(grid, rng, 6, base_period), 0.0, 1.0).astype(np.float32)
    col = np.array([[rng.uniform(0.3, 1.0) for _ in range(3)] for _ in range(8)], np.float32)
    op = (0.5 * np.arange(8) / 7.0).astype(np.float32)
    sp = 12.0 / (grid - 1)
    sc.volumes = [Volume(dens, col, op, origin=(-6.0, -6.0, -6.0), spacing=(sp, sp, sp))]
    sc.groups.append([])
    sc.group_volumes = [[], [0]]
    sc.instances.append(Instance(1))
    sc.name = "c4"
    return sc


# ---- OSPRay construction ---------------------------------------------------------------------

@dataclasses.dataclass
class OspScene:
    world: object
    camera: object
    renderer: object
    handles: list


def to_ospray(scene: Scene, renderer_params: Optional[dict] = None) -> OspScene:
    """Build the scene through the OSPRay C API (device must be current)."""
    from . import api as A
    from . import capi as K

    keep = []
    geo_h = []
    for g in scene.geometries:
        if g.type == "spheres":
            pos = A.shared_data(np.ascontiguousarray(g.spheres[:, :3], np.float32))
            rad = A.shared_data(np.ascontiguousarray(g.spheres[:, 3], np.float32))
            h = A.new("geometry", "sphere", {"sphere.position": pos, "sphere.radius": rad})
            keep += [pos, rad]
        else:
            params = {"vertex.position": A.shared_data(g.vertex), "index": A.shared_data(g.index)}
            if g.normal is not None:
                params["vertex.normal"] = A.shared_data(g.normal)
            h = A.new("geometry", "mesh", params)
            keep += list(params.values())
        geo_h.append(h)
    mat_h = []
    for m in scene.materials:
        p = {k: v for k, v in m.items() if k != "type"}
        if "ior" in p:  # metal: (wavelength, eta, k) rows as a vec3f array
            p["ior"] = A.shared_data(np.ascontiguousarray(p["ior"], np.float32).reshape(-1, 3))
            keep.append(p["ior"])
        mat_h.append(A.new("material", m["type"], p))
    model_h = []
    for m in scene.models:
        p = {"geometry": geo_h[m.geometry]}
        if m.material >= 0:
            p["material"] = mat_h[m.material]
        if m.user_id is not None:
            p["id"] = A.UInt(m.user_id)
        model_h.append(A.new("model", params=p))
    vmodel_h = []
    for v in scene.volumes:
        vol = A.new("volume", "structuredRegular",
                    {"data": A.shared_data3d(v.data), "gridOrigin": tuple(float(x) for x in v.origin),
                     "gridSpacing": tuple(float(x) for x in v.spacing)})
        tf = A.new("transfer_function", "piecewiseLinear",
                   {"color": A.shared_data(np.ascontiguousarray(v.tf_color, np.float32).reshape(-1, 3)),
                    "opacity": A.shared_data(np.ascontiguousarray(v.tf_opacity, np.float32).reshape(-1)),
                    "value": A.Box1f(*v.tf_range)})
        p = {"volume": vol, "transferFunction": tf, "densityScale": float(v.density_scale),
             "anisotropy": float(v.anisotropy)}
        if v.user_id is not None:
            p["id"] = A.UInt(v.user_id)
        vmodel_h.append(A.new("volumetric_model", params=p))
        keep += [vol, tf]
    group_h = []
    for gi, grp in enumerate(scene.groups):
        params = {}
        if grp:
            params["geometry"] = A.object_data([model_h[i] for i in grp], K.OSP_GEOMETRIC_MODEL)
        vols = scene.group_volumes[gi] if gi < len(scene.group_volumes) else []
        if vols:
            params["volume"] = A.object_data([vmodel_h[i] for i in vols], K.OSP_VOLUMETRIC_MODEL)
        group_h.append(A.new("group", params=params))
        keep += list(params.values())
    inst_h = []
    for inst in scene.instances:
        p = {"group": group_h[inst.group], "transform": A.Affine3f(inst.xfm)}
        if inst.user_id is not None:
            p["id"] = A.UInt(inst.user_id)
        inst_h.append(A.new("instance", params=p))
    light_h = []
    for l in scene.lights:
        p = {k: v for k, v in l.items() if k != "type"}
        light_h.append(A.new("light", l["type"], p))
    world_params = {}
    if inst_h:
        world_params["instance"] = A.object_data(inst_h, K.OSP_INSTANCE)
    if light_h:
        world_params["light"] = A.object_data(light_h, K.OSP_LIGHT)
    world = A.new("world", params=world_params)
    keep += list(world_params.values())
    cam = A.new("camera", "perspective", dict(scene.camera))
    rp = dict(scene.renderer)
    if renderer_params:
        rp.update(renderer_params)
    ren = A.new("renderer", "pathtracer", rp)
    handles = keep + geo_h + mat_h + model_h + vmodel_h + group_h + inst_h + light_h
    return OspScene(world, cam, ren, handles)

def near_tie_layers(width: int = 48, height: int = 48, instanced: bool = False) -> Scene:
    """Stacked near-coplanar surfaces for the near-tie re-resolution of
    closest hits: four 20x20 quads at z = 4 + k * 2^-21 (k = 0..3, one ulp
    of 4.0 apart), the farthest first in model order (so a float exact tie,
    which goes to the smaller key, picks the wrong layer), and in front of
    them two concentric spheres of radii 0.5 and 0.5 * (1 + 2^-22) at
    (0, 0, 3), the larger one first.  instanced: the group under a scale of
    0.5 and a small translation instead of the identity.  Camera at the
    origin looking along +z; Lambertian surfaces under an ambient light."""
    geoms = []
    for k in (3, 2, 1, 0):
        z = np.float32(4.0) + np.float32(k) * np.float32(2.0 ** -21)
        v = np.array([[-10, -10, z], [10, -10, z], [10, 10, z], [-10, 10, z]], np.float32)
        geoms.append(Geometry("quads", vertex=v, index=np.array([[0, 1, 2, 3]], np.uint32)))
    r0 = np.float32(0.5)
    r1 = np.float32(0.5) * (np.float32(1.0) + np.float32(2.0 ** -22))
    geoms.append(Geometry("spheres", spheres=np.array([[0.0, 0.0, 3.0, r1]], np.float32)))
    geoms.append(Geometry("spheres", spheres=np.array([[0.0, 0.0, 3.0, r0]], np.float32)))
    models = [Model(i, i % 2) for i in range(len(geoms))]
    insts = [Instance(0, xfm=(0.5, 0, 0, 0, 0.5, 0, 0, 0, 0.5, 0.1, -0.05, 0.0))] if instanced else [Instance(0)]
    mats = [{"type": "obj", "kd": (0.8, 0.3, 0.2)}, {"type": "obj", "kd": (0.2, 0.4, 0.8)}]
    cam = {"position": (0.0, 0.0, 0.0), "direction": (0.0, 0.0, 1.0), "up": (0.0, 1.0, 0.0), "fovy": 50.0}
    return Scene("near_tie_layers", width, height, geoms, models, [list(range(len(geoms)))], insts, mats,
                 [{"type": "ambient", "intensity": 1.0}], cam, {"pixelSamples": 2, "maxPathLength": 3})
