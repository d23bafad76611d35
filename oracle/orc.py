"""ORACLE — test infrastructure only.

ctypes front end of oracle/liborc.so, the CPU restatement of the reference
path tracer (see oracle/refcpu/*.h for the reference files each part
follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it, as the checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORC_LIB names another build of the oracle (for instance one with a planted
# error, to show that the pinning tests reject it)
LIBPATH = os.environ.get("ORC_LIB", os.path.join(HERE, "liborc.so"))

ORC_GEOM = {"triangles": 0, "quads": 1, "spheres": 2}
ORC_MAT = {"obj": 0, "metal": 1, "glass": 2, "luminous": 3}
ORC_LIGHT = {"quad": 0, "distant": 1, "ambient": 2}


class OrcGeometry(C.Structure):
    _fields_ = [("type", C.c_int32), ("vertex", C.c_void_p), ("numVertices", C.c_int64), ("index", C.c_void_p),
                ("numPrimitives", C.c_int64), ("normal", C.c_void_p), ("spheres", C.c_void_p)]


class OrcModel(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("material", C.c_int32), ("userID", C.c_uint32), ("invertNormals", C.c_int32)]


class OrcGroup(C.Structure):
    _fields_ = [("firstModel", C.c_int32), ("numModels", C.c_int32), ("firstVolume", C.c_int32),
                ("numVolumes", C.c_int32)]


class OrcVolume(C.Structure):
    _fields_ = [("voxels", C.c_void_p), ("dims", C.c_int64 * 3), ("gridOrigin", C.c_float * 3),
                ("gridSpacing", C.c_float * 3), ("tfColor", C.c_void_p), ("numColors", C.c_int32),
                ("tfOpacity", C.c_void_p), ("numOpacities", C.c_int32), ("tfValueRange", C.c_float * 2),
                ("densityScale", C.c_float), ("anisotropy", C.c_float), ("userID", C.c_uint32)]


class OrcInstance(C.Structure):
    _fields_ = [("group", C.c_int32), ("userID", C.c_uint32), ("xfm", C.c_float * 12)]


class OrcMaterial(C.Structure):
    _fields_ = [("type", C.c_int32), ("kd", C.c_float * 3), ("d", C.c_float), ("tf", C.c_float * 3),
                ("eta", C.c_float * 3), ("k", C.c_float * 3), ("roughness", C.c_float), ("etaInside", C.c_float),
                ("etaOutside", C.c_float), ("attenuationColorInside", C.c_float * 3),
                ("attenuationColorOutside", C.c_float * 3), ("attenuationDistance", C.c_float),
                ("color", C.c_float * 3), ("intensity", C.c_float), ("transparency", C.c_float),
                ("ks", C.c_float * 3), ("ns", C.c_float), ("ior", C.c_void_p), ("numIor", C.c_int64)]


class OrcLight(C.Structure):
    _fields_ = [("type", C.c_int32), ("visible", C.c_int32), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("intensityQuantity", C.c_int32), ("position", C.c_float * 3), ("edge1", C.c_float * 3),
                ("edge2", C.c_float * 3), ("direction", C.c_float * 3), ("angularDiameter", C.c_float)]


class OrcScene(C.Structure):
    _fields_ = [("geometries", C.POINTER(OrcGeometry)), ("numGeometries", C.c_int32),
                ("models", C.POINTER(OrcModel)), ("numModels", C.c_int32),
                ("groupModels", C.POINTER(C.c_int32)), ("groups", C.POINTER(OrcGroup)), ("numGroups", C.c_int32),
                ("instances", C.POINTER(OrcInstance)), ("numInstances", C.c_int32),
                ("materials", C.POINTER(OrcMaterial)), ("numMaterials", C.c_int32),
                ("lights", C.POINTER(OrcLight)), ("numLights", C.c_int32),
                ("volumes", C.POINTER(OrcVolume)), ("numVolumes", C.c_int32),
                ("groupVolumes", C.POINTER(C.c_int32))]


class OrcCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("aspect", C.c_float), ("nearClip", C.c_float)]


class OrcRenderer(C.Structure):
    _fields_ = [("pixelSamples", C.c_int32), ("maxPathLength", C.c_int32), ("roulettePathLength", C.c_int32),
                ("maxScatteringEvents", C.c_int32), ("lightSamples", C.c_int32),
                ("limitIndirectLightSamples", C.c_int32), ("minContribution", C.c_float),
                ("maxContribution", C.c_float), ("backgroundColor", C.c_float * 4), ("pixelFilter", C.c_int32),
                ("varianceThreshold", C.c_float), ("geometryLights", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIBPATH)
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(OrcScene)]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_fb_create.restype = C.c_void_p
        L.orc_fb_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32]
        L.orc_fb_destroy.argtypes = [C.c_void_p]
        L.orc_fb_reset.argtypes = [C.c_void_p]
        L.orc_fb_set_projected_depth.argtypes = [C.c_void_p, C.c_int]
        L.orc_render_frame.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(OrcCamera), C.POINTER(OrcRenderer),
                                       C.c_int, C.c_int, C.c_int]
        L.orc_fb_variance.restype = C.c_float
        L.orc_fb_variance.argtypes = [C.c_void_p, C.POINTER(OrcRenderer)]
        L.orc_fb_map.restype = C.c_void_p
        L.orc_fb_map.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_fb_accum_color.restype = C.c_void_p
        L.orc_fb_accum_color.argtypes = [C.c_void_p]
        L.orc_fb_task_variance.restype = C.c_void_p
        L.orc_fb_task_variance.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_sobol_rev_table.argtypes = [C.c_void_p]
        L.orc_hash.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_pcg_raw.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p, C.c_int]
        L.orc_pcg_init_sequence.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_int]
        L.orc_ld_samples5.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        L.orc_fmath.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_pixel_filter.argtypes = [C.POINTER(OrcRenderer), C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_camera_rays.argtypes = [C.POINTER(OrcCamera), C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_intersect.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_uint64)]
        L.orc_volume_sample.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_tf_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_tf_max_opacity.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_volume_intervals.restype = C.c_int
        L.orc_volume_intervals.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_volume_estimates.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_uint32, C.c_void_p,
                                           C.c_void_p, C.c_void_p]
        L.orc_hg_sample.argtypes = [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int64]
        L.orc_hg_eval.argtypes = [C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_variance_buffers.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                                           C.c_void_p]
        L.orc_distribution1d.restype = C.c_float
        L.orc_distribution1d.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.orc_light_samples.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4 + [C.c_int64, C.c_void_p]
        L.orc_select_light.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _f3(v, default=(0.0, 0.0, 0.0)):
    v = default if v is None else v
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def make_camera(cam: dict, near_clip: float = 1e-6) -> OrcCamera:
    c = OrcCamera()
    c.position = _f3(cam.get("position"))
    c.direction = _f3(cam.get("direction"), (0, 0, 1))
    c.up = _f3(cam.get("up"), (0, 1, 0))
    c.fovy = cam.get("fovy", 60.0)
    c.aspect = cam.get("aspect", 1.0)
    c.nearClip = cam.get("nearClip", near_clip)
    return c


def make_renderer(p: dict) -> OrcRenderer:
    """Renderer parameters with the reference defaults (Renderer.cpp:38-92,
    PathTracer.cpp:29-51)."""
    r = OrcRenderer()
    r.pixelSamples = int(p.get("pixelSamples", 1))
    r.maxPathLength = int(p.get("maxPathLength", 20))
    r.roulettePathLength = int(p.get("roulettePathLength", 5))
    r.maxScatteringEvents = int(p.get("maxScatteringEvents", 20))
    r.lightSamples = int(p.get("lightSamples", -1))
    r.limitIndirectLightSamples = int(bool(p.get("limitIndirectLightSamples", True)))
    r.minContribution = float(p.get("minContribution", 0.001))
    r.maxContribution = float(p.get("maxContribution", float("inf")))
    bg = p.get("backgroundColor", (0.0, 0.0, 0.0, 0.0))
    if isinstance(bg, (int, float)):
        bg = (bg, bg, bg, 0.0)
    if len(bg) == 3:
        bg = tuple(bg) + (0.0,)
    r.backgroundColor = (C.c_float * 4)(*bg)
    pf = p.get("pixelFilter", 2)
    r.pixelFilter = int(getattr(pf, "v", pf))
    r.varianceThreshold = float(p.get("varianceThreshold", 0.0))
    r.geometryLights = int(bool(p.get("geometryLights", True)))
    return r


class OracleScene:
    """The Scene packed for liborc.so, with its BVHs built."""

    def __init__(self, scene):
        L = lib()
        self._keep = []
        geos = (OrcGeometry * len(scene.geometries))()
        for i, g in enumerate(scene.geometries):
            geos[i].type = ORC_GEOM[g.type]
            if g.type == "spheres":
                sp = np.ascontiguousarray(g.spheres, np.float32)
                self._keep.append(sp)
                geos[i].spheres = sp.ctypes.data
                geos[i].numPrimitives = sp.shape[0]
            else:
                v = np.ascontiguousarray(g.vertex, np.float32)
                ix = np.ascontiguousarray(g.index, np.uint32)
                self._keep += [v, ix]
                geos[i].vertex = v.ctypes.data
                geos[i].numVertices = v.shape[0]
                geos[i].index = ix.ctypes.data
                geos[i].numPrimitives = ix.shape[0]
                if g.normal is not None:
                    nrm = np.ascontiguousarray(g.normal, np.float32)
                    self._keep.append(nrm)
                    geos[i].normal = nrm.ctypes.data
        models = (OrcModel * len(scene.models))()
        for i, m in enumerate(scene.models):
            models[i].geometry = m.geometry
            models[i].material = m.material
            models[i].userID = 0xFFFFFFFF if m.user_id is None else m.user_id
        gm, gv, groups = [], [], (OrcGroup * len(scene.groups))()
        group_volumes = getattr(scene, "group_volumes", None) or []
        for i, grp in enumerate(scene.groups):
            groups[i].firstModel = len(gm)
            groups[i].numModels = len(grp)
            gm += list(grp)
            vols = group_volumes[i] if i < len(group_volumes) else []
            groups[i].firstVolume = len(gv)
            groups[i].numVolumes = len(vols)
            gv += list(vols)
        gmodels = (C.c_int32 * max(1, len(gm)))(*gm)
        gvols = (C.c_int32 * max(1, len(gv)))(*gv)
        volumes = getattr(scene, "volumes", None) or []
        vols = (OrcVolume * max(1, len(volumes)))()
        for i, v in enumerate(volumes):
            data = np.ascontiguousarray(v.data, np.float32)  # (nz, ny, nx)
            col = np.ascontiguousarray(v.tf_color, np.float32).reshape(-1, 3)
            op = np.ascontiguousarray(v.tf_opacity, np.float32).reshape(-1)
            self._keep += [data, col, op]
            vols[i].voxels = data.ctypes.data
            vols[i].dims = (C.c_int64 * 3)(data.shape[2], data.shape[1], data.shape[0])
            vols[i].gridOrigin = _f3(v.origin)
            vols[i].gridSpacing = _f3(v.spacing, (1, 1, 1))
            vols[i].tfColor = col.ctypes.data
            vols[i].numColors = col.shape[0]
            vols[i].tfOpacity = op.ctypes.data
            vols[i].numOpacities = op.shape[0]
            vols[i].tfValueRange = (C.c_float * 2)(*[float(np.float32(x)) for x in v.tf_range])
            vols[i].densityScale = v.density_scale
            vols[i].anisotropy = v.anisotropy
            vols[i].userID = 0xFFFFFFFF if v.user_id is None else v.user_id
        insts = (OrcInstance * len(scene.instances))()
        for i, inst in enumerate(scene.instances):
            insts[i].group = inst.group
            insts[i].userID = 0xFFFFFFFF if inst.user_id is None else inst.user_id
            insts[i].xfm = (C.c_float * 12)(*[float(np.float32(x)) for x in inst.xfm])
        mats = (OrcMaterial * max(1, len(scene.materials)))()
        for i, m in enumerate(scene.materials):
            t = m["type"]
            mats[i].type = ORC_MAT[t]
            mats[i].kd = _f3(m.get("kd"), (0.8, 0.8, 0.8))
            mats[i].d = m.get("d", 1.0)
            mats[i].tf = _f3(m.get("tf"))
            mats[i].eta = _f3(m.get("eta") if t == "metal" else None, (1.47, 0.984, 0.602))
            mats[i].k = _f3(m.get("k"), (7.64, 6.55, 5.36))
            mats[i].roughness = m.get("roughness", 0.1)
            eta = m.get("eta", 1.5) if t == "glass" else 1.5
            mats[i].etaInside = m.get("etaInside", eta)
            mats[i].etaOutside = m.get("etaOutside", 1.0)
            ac = m.get("attenuationColor", (1.0, 1.0, 1.0))
            mats[i].attenuationColorInside = _f3(m.get("attenuationColorInside", ac))
            mats[i].attenuationColorOutside = _f3(m.get("attenuationColorOutside", (1.0, 1.0, 1.0)))
            mats[i].attenuationDistance = m.get("attenuationDistance", 1.0)
            mats[i].color = _f3(m.get("color"), (1.0, 1.0, 1.0))  # Luminous.cpp:37-48
            mats[i].intensity = m.get("intensity", 1.0)
            mats[i].transparency = m.get("transparency", 0.0)
            mats[i].ks = _f3(m.get("ks"))  # OBJ.cpp:32-37
            mats[i].ns = m.get("ns", 10.0)
            if t == "metal" and m.get("ior") is not None:  # Metal.cpp:30-57
                ior = np.ascontiguousarray(m["ior"], np.float32).reshape(-1, 3)
                self._keep.append(ior)
                mats[i].ior = ior.ctypes.data
                mats[i].numIor = len(ior)
        lights = (OrcLight * max(1, len(scene.lights)))()
        for i, l in enumerate(scene.lights):
            lights[i].type = ORC_LIGHT[l["type"]]
            lights[i].visible = int(l.get("visible", True))
            lights[i].color = _f3(l.get("color"), (1, 1, 1))
            lights[i].intensity = l.get("intensity", 1.0)
            iq = l.get("intensityQuantity", -1)
            lights[i].intensityQuantity = int(getattr(iq, "v", iq))
            lights[i].position = _f3(l.get("position"))
            lights[i].edge1 = _f3(l.get("edge1"), (1, 0, 0))
            lights[i].edge2 = _f3(l.get("edge2"), (0, 1, 0))
            lights[i].direction = _f3(l.get("direction"), (0, 0, 1))
            lights[i].angularDiameter = l.get("angularDiameter", 0.0)
        self._keep += [geos, models, groups, gmodels, insts, mats, lights, gvols, vols]
        desc = OrcScene(geos, len(scene.geometries), models, len(scene.models), gmodels, groups, len(scene.groups),
                        insts, len(scene.instances), mats, len(scene.materials), lights, len(scene.lights),
                        vols, len(volumes), gvols)
        self._desc = desc
        self.h = L.orc_scene_create(C.byref(desc))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_scene_destroy(self.h)
            self.h = None

    def intersect(self, rays8: np.ndarray, fp64: bool = False, float_order: bool = False):
        """Closest hits: the extension-ray traversal (near-ties re-resolved
        in FP64); fp64: the FP64 brute force; float_order: the traversal
        with (float t, key) order only."""
        L = lib()
        r = np.ascontiguousarray(rays8, np.float32)
        n = r.shape[0]
        tuv = np.empty((n, 3), np.float32)
        ids = np.empty((n, 3), np.uint32)
        st = (C.c_uint64 * 2)()
        L.orc_intersect(self.h, r.ctypes.data, n, 1 if fp64 else (2 if float_order else 0), tuv.ctypes.data,
                        ids.ctypes.data, st)
        return tuv, ids, (int(st[0]), int(st[1]))


    def light_samples(self, light: int, P, Ng, Ns, s2):
        """lightSample of light `light` (index into [geometry lights, world
        lights]) at P: columns dir (3), dist, weight (3), pdf."""
        s2 = np.ascontiguousarray(s2, np.float32).reshape(-1, 2)
        vecs = [np.ascontiguousarray(v, np.float32).reshape(3) for v in (P, Ng, Ns)]
        out = np.empty((s2.shape[0], 8), np.float32)
        lib().orc_light_samples(self.h, light, vecs[0].ctypes.data, vecs[1].ctypes.data, vecs[2].ctypes.data,
                                s2.ctypes.data, s2.shape[0], out.ctypes.data)
        return out

    def select_light(self, ss, geometry_lights: bool = True):
        """The path tracer's light selection for each value of ss: (index
        into the light list, selection probability)."""
        ss = np.ascontiguousarray(ss, np.float32).reshape(-1)
        idx = np.empty(ss.size, np.int32)
        prob = np.empty(ss.size, np.float32)
        lib().orc_select_light(self.h, int(geometry_lights), ss.ctypes.data, ss.size, idx.ctypes.data,
                               prob.ctypes.data)
        return idx, prob

    # ---- volume hooks ----
    def volume_sample(self, vol: int, pts: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.float32)
        lib().orc_volume_sample(self.h, vol, p.ctypes.data, out.ctypes.data, p.shape[0])
        return out

    def tf_get(self, vol: int, values) -> np.ndarray:
        v = np.ascontiguousarray(values, np.float32).reshape(-1)
        out = np.empty((v.size, 4), np.float32)
        lib().orc_tf_get(self.h, vol, v.ctypes.data, out.ctypes.data, v.size)
        return out

    def tf_max_opacity(self, vol: int, ranges) -> np.ndarray:
        r = np.ascontiguousarray(ranges, np.float32).reshape(-1, 2)
        out = np.empty(r.shape[0], np.float32)
        lib().orc_tf_max_opacity(self.h, vol, r.ctypes.data, out.ctypes.data, r.shape[0])
        return out

    def volume_intervals(self, vol: int, ray8) -> np.ndarray:
        r = np.ascontiguousarray(ray8, np.float32).reshape(8)
        out = np.empty((4096, 4), np.float32)
        n = lib().orc_volume_intervals(self.h, vol, r.ctypes.data, out.ctypes.data, 4096)
        assert n <= 4096
        return out[:n].copy()

    def volume_estimates(self, rays8, trials: int, seed: int = 7):
        """Per world ray: (mean transmittance, fraction of free-path
        collisions, mean collision distance) over `trials` PCG streams."""
        r = np.ascontiguousarray(rays8, np.float32).reshape(-1, 8)
        n = r.shape[0]
        T, hf, md = (np.empty(n, np.float32) for _ in range(3))
        lib().orc_volume_estimates(self.h, r.ctypes.data, n, trials, seed, T.ctypes.data, hf.ctypes.data,
                                   md.ctypes.data)
        return T, hf, md


def hg_sample(g: float, wo, s2):
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    s2 = np.ascontiguousarray(s2, np.float32).reshape(-1, 2)
    n = s2.shape[0]
    wo = np.broadcast_to(wo, (n, 3)).copy()
    wi = np.empty((n, 3), np.float32)
    pdf = np.empty(n, np.float32)
    typ = np.empty(n, np.uint32)
    lib().orc_hg_sample(g, wo.ctypes.data, s2.ctypes.data, wi.ctypes.data, pdf.ctypes.data, typ.ctypes.data, n)
    return wi, pdf, typ


def hg_eval(g: float, wo, wi):
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    n = wi.shape[0]
    wo = np.broadcast_to(np.ascontiguousarray(wo, np.float32).reshape(-1, 3), (n, 3)).copy()
    out = np.empty(n, np.float32)
    lib().orc_hg_eval(g, wo.ctypes.data, wi.ctypes.data, out.ctypes.data, n)
    return out


class OracleFrameBuffer:
    def __init__(self, w: int, h: int, fmt: int, channels: int):
        self.w, self.h, self.fmt = w, h, fmt
        self.ptr = lib().orc_fb_create(w, h, fmt, channels)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_fb_destroy(self.ptr)
            self.ptr = None

    def render(self, scene: OracleScene, camera: OrcCamera, renderer: OrcRenderer, nthreads: int = 8,
               rows: tuple = (0, -1)):
        lib().orc_render_frame(scene.h, self.ptr, C.byref(camera), C.byref(renderer), nthreads, rows[0], rows[1])

    def variance(self, renderer: OrcRenderer) -> float:
        return float(lib().orc_fb_variance(self.ptr, C.byref(renderer)))

    def reset(self):
        lib().orc_fb_reset(self.ptr)

    def channel(self, ch: int) -> np.ndarray:
        p = lib().orc_fb_map(self.ptr, ch)
        if not p:
            raise ValueError(f"channel {ch} not present")
        n = self.w * self.h
        if ch == 1:
            if self.fmt == 3:
                return np.ctypeslib.as_array((C.c_float * (4 * n)).from_address(p)).copy().reshape(self.h, self.w, 4)
            return np.ctypeslib.as_array((C.c_uint8 * (4 * n)).from_address(p)).copy().reshape(self.h, self.w, 4)
        if ch == 2:
            return np.ctypeslib.as_array((C.c_float * n).from_address(p)).copy().reshape(self.h, self.w)
        if ch in (64, 128, 256):
            return np.ctypeslib.as_array((C.c_uint32 * n).from_address(p)).copy().reshape(self.h, self.w)
        return np.ctypeslib.as_array((C.c_float * (3 * n)).from_address(p)).copy().reshape(self.h, self.w, 3)

    def task_variance(self) -> np.ndarray:
        """Per-8x8-task error of the last variance pass (row-major tasks)."""
        n = C.c_int(0)
        p = lib().orc_fb_task_variance(self.ptr, C.byref(n))
        return np.ctypeslib.as_array((C.c_float * n.value).from_address(p)).copy()

    def accum_color(self) -> np.ndarray:
        p = lib().orc_fb_accum_color(self.ptr)
        n = self.w * self.h
        return np.ctypeslib.as_array((C.c_float * (4 * n)).from_address(p)).copy().reshape(self.h, self.w, 4)


def render_scene(scene, frames: int = 1, fmt: int = 3, channels: int = 1 | 4, nthreads: int = 8,
                 renderer_params: dict | None = None, oscene: OracleScene | None = None, rows=(0, -1)):
    """Render `frames` accumulated frames of a Scene with the oracle."""
    os_ = oscene or OracleScene(scene)
    fb = OracleFrameBuffer(scene.width, scene.height, fmt, channels)
    rp = dict(scene.renderer)
    if renderer_params:
        rp.update(renderer_params)
    cam = make_camera(scene.camera)
    ren = make_renderer(rp)
    for _ in range(frames):
        fb.render(os_, cam, ren, nthreads, rows)
    return fb, os_, ren


def variance_buffers(color4: np.ndarray, variance4: np.ndarray, threshold: float):
    """Variance_kernel + getAvgError on given (h, w, 4) accumulated colour
    and variance buffers: (per-task errors (ceil(h/8), ceil(w/8)), frame
    estimate)."""
    c = np.ascontiguousarray(color4, np.float32)
    v = np.ascontiguousarray(variance4, np.float32)
    h, w = c.shape[:2]
    tasks = np.empty(((h + 7) // 8, (w + 7) // 8), np.float32)
    avg = C.c_float()
    lib().orc_variance_buffers(w, h, c.ctypes.data, v.ctypes.data, threshold, tasks.ctypes.data, C.byref(avg))
    return tasks, float(avg.value)


def distribution1d(weights, s):
    """Distribution1D_create(weights) then _sample(s): (sum, idx, prob, pdf,
    frac)."""
    w = np.ascontiguousarray(weights, np.float32).reshape(-1)
    s = np.ascontiguousarray(s, np.float32).reshape(-1)
    idx = np.empty(s.size, np.int32)
    out = np.empty((s.size, 3), np.float32)
    total = lib().orc_distribution1d(w.ctypes.data, w.size, s.ctypes.data, s.size, idx.ctypes.data, out.ctypes.data)
    return float(total), idx, out[:, 0], out[:, 1], out[:, 2]


def pixel_filter(renderer: dict, s2):
    """Pixel-filter offsets of the renderer's filter for sample pairs s2."""
    s2 = np.ascontiguousarray(s2, np.float32).reshape(-1, 2)
    out = np.empty_like(s2)
    r = make_renderer(renderer)
    lib().orc_pixel_filter(C.byref(r), s2.ctypes.data, out.ctypes.data, s2.shape[0])
    return out


def camera_rays(camera: dict, screen2, near_clip: float = 1e-6):
    """Camera rays (org (3), dir (3)) for screen positions in [0,1]^2."""
    sc = np.ascontiguousarray(screen2, np.float32).reshape(-1, 2)
    out = np.empty((sc.shape[0], 6), np.float32)
    cam = make_camera(camera, near_clip)
    lib().orc_camera_rays(C.byref(cam), sc.ctypes.data, out.ctypes.data, sc.shape[0])
    return out
