"""Small Pythonic layer over the C API bindings (ospray_amd.capi).

Parameter helpers infer the OSPDataType from the Python value, array helpers
wrap numpy arrays with ospNewSharedData (the arrays are kept alive by the
returned handle objects), and errors raised by the device are surfaced as
Python exceptions (the C API itself reports them through ospDeviceGet-
LastErrorCode, ospray/api/API.cpp:45-66).
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np

from . import capi as K


class OSPRayError(RuntimeError):
    pass


_keepalive: list = []


def _b(s: str) -> bytes:
    return s.encode()


def init(device: str = "hip", **params) -> int:
    """ospLoadModule + ospNewDevice + ospDeviceCommit + ospSetCurrentDevice."""
    L = K.lib()
    err = L.ospLoadModule(_b(device))
    if err != K.OSP_NO_ERROR:
        raise OSPRayError(f"ospLoadModule('{device}') failed with error {err}")
    dev = L.ospNewDevice(_b(device))
    if not dev:
        raise OSPRayError(f"ospNewDevice('{device}') failed")
    for k, v in params.items():
        set_device_param(dev, k, v)
    L.ospDeviceSetErrorCallback(dev, _on_error, None)
    L.ospDeviceCommit(dev)
    L.ospSetCurrentDevice(dev)
    check()
    return dev


def set_device_param(dev: int, name: str, v) -> None:
    """ospDeviceSetParam for bool, int and string values."""
    L = K.lib()
    if isinstance(v, bool):
        x = C.c_int(int(v))
        L.ospDeviceSetParam(dev, _b(name), K.OSP_BOOL, C.byref(x))
    elif isinstance(v, int):
        x = C.c_int(v)
        L.ospDeviceSetParam(dev, _b(name), K.OSP_INT, C.byref(x))
    elif isinstance(v, str):
        L.ospDeviceSetParam(dev, _b(name), K.OSP_STRING, C.c_char_p(_b(v)))
    else:
        raise TypeError(f"device parameter {name}: unsupported type")


_errors: list = []


@K.ERROR_CALLBACK
def _on_error(user, code, msg):
    _errors.append((int(code), msg.decode() if msg else ""))


def check() -> None:
    """Raise if the device reported an error (through its error callback)
    since the last check."""
    if _errors:
        code, msg = _errors[0]
        _errors.clear()
        raise OSPRayError(f"OSPRay error {code}: {msg}")


def set_param(obj: int, name: str, value) -> None:
    L = K.lib()
    n = _b(name)
    if isinstance(value, bool):
        L.ospSetBool(obj, n, int(value))
    elif isinstance(value, (int, np.integer)) and not isinstance(value, bool):
        L.ospSetInt(obj, n, int(value))
    elif isinstance(value, (float, np.floating)):
        L.ospSetFloat(obj, n, float(value))
    elif isinstance(value, str):
        L.ospSetString(obj, n, _b(value))
    elif isinstance(value, UInt):
        L.ospSetUInt(obj, n, value.v)
    elif isinstance(value, Affine3f):
        arr = (C.c_float * 12)(*value.v)
        L.ospSetParam(obj, n, K.OSP_AFFINE3F, arr)
    elif isinstance(value, Box1f):
        arr = (C.c_float * 2)(*value.v)
        L.ospSetParam(obj, n, K.OSP_BOX1F, arr)
    elif isinstance(value, Handle):
        h = C.c_void_p(value.h)
        L.ospSetParam(obj, n, value.type, C.byref(h))
    elif isinstance(value, (tuple, list, np.ndarray)) and len(value) in (2, 3, 4):
        vals = [float(x) for x in value]
        if len(vals) == 2:
            L.ospSetVec2f(obj, n, *vals)
        elif len(vals) == 3:
            L.ospSetVec3f(obj, n, *vals)
        else:
            L.ospSetVec4f(obj, n, *vals)
    else:
        raise TypeError(f"parameter {name}: unsupported value {value!r}")


def set_params(obj: int, params: dict) -> None:
    for k, v in params.items():
        set_param(obj, k, v)


class UInt:
    """Marks an integer parameter as OSP_UINT (e.g. 'id', 'pixelFilter')."""

    def __init__(self, v: int):
        self.v = int(v)


class Box1f:
    """box1f parameter (e.g. a transfer function's 'value' range)."""

    def __init__(self, lo: float, hi: float):
        self.v = [float(lo), float(hi)]


class Affine3f:
    """affine3f parameter: linear columns vx, vy, vz then translation p."""

    def __init__(self, v: Sequence[float]):
        assert len(v) == 12
        self.v = [float(x) for x in v]


class Handle:
    """An OSPRay object handle with its OSPDataType (and owned references)."""

    def __init__(self, h: int, type_: int, keep=None):
        if not h:
            check()
            raise OSPRayError("object creation failed")
        self.h = h
        self.type = type_
        self.keep = keep

    def release(self):
        if self.h:
            K.lib().ospRelease(self.h)
            self.h = 0


_NP_TYPES = {
    (np.dtype(np.float32), 3): K.OSP_VEC3F,
    (np.dtype(np.float32), 1): K.OSP_FLOAT,
    (np.dtype(np.float32), 4): K.OSP_VEC4F,
    (np.dtype(np.uint32), 3): K.OSP_VEC3UI,
    (np.dtype(np.uint32), 4): K.OSP_VEC4UI,
    (np.dtype(np.uint32), 1): K.OSP_UINT,
}


def shared_data(arr: np.ndarray) -> Handle:
    """ospNewSharedData over a contiguous numpy array of float32/uint32
    rows (1, 3 or 4 components); the array is kept alive by the handle."""
    a = np.ascontiguousarray(arr)
    comps = 1 if a.ndim == 1 else a.shape[1]
    t = _NP_TYPES.get((a.dtype, comps))
    if t is None:
        raise TypeError(f"no OSPDataType for {a.dtype} x {comps}")
    n = a.shape[0]
    h = K.lib().ospNewSharedData(a.ctypes.data, t, n, 0, 1, 0, 1, 0, None, None)
    return Handle(h, K.OSP_DATA, keep=a)


def shared_data3d(arr: np.ndarray) -> Handle:
    """ospNewSharedData over a C-contiguous (nz, ny, nx) float32 array: a 3D
    array with x fastest, as structured volumes take it."""
    a = np.ascontiguousarray(arr, np.float32)
    if a.ndim != 3:
        raise ValueError("expected a 3D array")
    nz, ny, nx = a.shape
    h = K.lib().ospNewSharedData(a.ctypes.data, K.OSP_FLOAT, nx, 0, ny, 0, nz, 0, None, None)
    return Handle(h, K.OSP_DATA, keep=a)


def object_data(handles: Iterable[Handle], type_: int) -> Handle:
    """A data array of object handles (copied with ospCopyData, so the
    array holds its own references)."""
    hs = list(handles)
    arr = (C.c_void_p * max(1, len(hs)))(*[x.h for x in hs])
    L = K.lib()
    tmp = L.ospNewSharedData(arr, type_, len(hs), 0, 1, 0, 1, 0, None, None)
    data = L.ospNewData(type_, len(hs), 1, 1)
    L.ospCopyData(tmp, data, 0, 0, 0)
    L.ospRelease(tmp)
    return Handle(data, K.OSP_DATA)


def new(kind: str, type_: str = "", params: dict | None = None, commit: bool = True) -> Handle:
    L = K.lib()
    ctor = {
        "camera": (L.ospNewCamera, K.OSP_CAMERA),
        "geometry": (L.ospNewGeometry, K.OSP_GEOMETRY),
        "material": (L.ospNewMaterial, K.OSP_MATERIAL),
        "light": (L.ospNewLight, K.OSP_LIGHT),
        "renderer": (L.ospNewRenderer, K.OSP_RENDERER),
        "volume": (L.ospNewVolume, K.OSP_VOLUME),
        "texture": (L.ospNewTexture, K.OSP_TEXTURE),
    }
    if kind in ctor:
        fn, t = ctor[kind]
        h = Handle(fn(_b(type_)), t)
    elif kind == "model":
        h = Handle(L.ospNewGeometricModel(None), K.OSP_GEOMETRIC_MODEL)
    elif kind == "transfer_function":
        h = Handle(L.ospNewTransferFunction(_b(type_)), K.OSP_TRANSFER_FUNCTION)
    elif kind == "volumetric_model":
        h = Handle(L.ospNewVolumetricModel(None), K.OSP_VOLUMETRIC_MODEL)
    elif kind == "group":
        h = Handle(L.ospNewGroup(), K.OSP_GROUP)
    elif kind == "instance":
        h = Handle(L.ospNewInstance(None), K.OSP_INSTANCE)
    elif kind == "world":
        h = Handle(L.ospNewWorld(), K.OSP_WORLD)
    else:
        raise ValueError(kind)
    if params:
        set_params(h.h, params)
    if commit:
        L.ospCommit(h.h)
        check()
    return h


def commit(h: Handle) -> None:
    K.lib().ospCommit(h.h)
    check()


def framebuffer(w: int, h: int, fmt: int = K.OSP_FB_RGBA32F,
                channels: int = K.OSP_FB_COLOR | K.OSP_FB_ACCUM, params: dict | None = None) -> Handle:
    fb = Handle(K.lib().ospNewFrameBuffer(w, h, fmt, channels), K.OSP_FRAMEBUFFER)
    if params:
        set_params(fb.h, params)
    K.lib().ospCommit(fb.h)
    check()
    return fb


def render(fb: Handle, renderer: Handle, camera: Handle, world: Handle) -> float:
    """Render one frame, wait, return the task duration in seconds."""
    L = K.lib()
    f = L.ospRenderFrame(fb.h, renderer.h, camera.h, world.h)
    if not f:
        check()
        raise OSPRayError("ospRenderFrame failed")
    L.ospWait(f, K.OSP_TASK_FINISHED)
    d = L.ospGetTaskDuration(f)
    L.ospRelease(f)
    check()
    return d


def map_channel(fb: Handle, channel: int, w: int, h: int) -> np.ndarray:
    """Copy of a framebuffer channel as a numpy array (ospMapFrameBuffer)."""
    L = K.lib()
    p = L.ospMapFrameBuffer(fb.h, channel)
    if not p:
        check()
        raise OSPRayError(f"channel {channel} not available")
    n = w * h
    if channel == K.OSP_FB_COLOR:
        raise ValueError("use map_color")
    if channel in (K.OSP_FB_DEPTH,):
        out = np.ctypeslib.as_array((C.c_float * n).from_address(p)).copy()
    elif channel in (K.OSP_FB_ID_PRIMITIVE, K.OSP_FB_ID_OBJECT, K.OSP_FB_ID_INSTANCE):
        out = np.ctypeslib.as_array((C.c_uint32 * n).from_address(p)).copy()
    else:
        out = np.ctypeslib.as_array((C.c_float * (3 * n)).from_address(p)).copy().reshape(n, 3)
    L.ospUnmapFrameBuffer(p, fb.h)
    return out.reshape((h, w) + out.shape[1:]) if out.ndim == 1 else out.reshape(h, w, 3)


def map_color(fb: Handle, w: int, h: int, fmt: int) -> np.ndarray:
    L = K.lib()
    p = L.ospMapFrameBuffer(fb.h, K.OSP_FB_COLOR)
    if not p:
        check()
        raise OSPRayError("colour channel not available")
    n = w * h
    if fmt == K.OSP_FB_RGBA32F:
        out = np.ctypeslib.as_array((C.c_float * (4 * n)).from_address(p)).copy().reshape(h, w, 4)
    else:
        out = np.ctypeslib.as_array((C.c_uint8 * (4 * n)).from_address(p)).copy().reshape(h, w, 4)
    L.ospUnmapFrameBuffer(p, fb.h)
    return out


def kernel_stats() -> dict:
    """{stage: (launches, total ms)} from ospHipGetKernelStats (per-bounce
    timers: up to 16 stages of each of up to 64 bounces)."""
    cap = 1024
    arr = (K.OSPHipKernelStat * cap)()
    n = K.ext().ospHipGetKernelStats(arr, cap)
    return {arr[i].name.decode(): (arr[i].launches, arr[i].totalMs) for i in range(max(n, 0))}


def reset_kernel_stats() -> None:
    K.ext().ospHipResetKernelStats()


def traversal_stats() -> tuple:
    """(node visits, primitive tests, traversals) of the extension rays, of
    the shadow-ray occlusion pass and of the shadow-ray transparency pass, of
    the last frame rendered with "traversalStats"."""
    out = (C.c_uint64 * 9)()
    K.ext().ospHipGetTraversalStats(out)
    return tuple(int(x) for x in out)


def nrc_state(renderer: Handle) -> tuple:
    """(loss of the last training step, Adam steps so far, cache queries and
    training records of the last frame) of a renderer's radiance cache."""
    loss, steps, q, t = C.c_float(0.0), C.c_int(0), C.c_int64(0), C.c_int64(0)
    if K.ext().ospHipGetNrcState(renderer.h, C.byref(loss), C.byref(steps), C.byref(q), C.byref(t)) != 0:
        check()
        raise OSPRayError("ospHipGetNrcState failed")
    return float(loss.value), int(steps.value), int(q.value), int(t.value)


def _handles(hs):
    return (C.c_void_p * len(hs))(*[h.h for h in hs])


def render_batch(fbs, renderers, cameras, world: Handle, graph: bool = True):
    """Render a batch of views in one enqueue (a HIP graph when graph is
    True); returns each view's GPU milliseconds from the batch start."""
    n = len(fbs)
    ms = (C.c_float * n)()
    rc = K.ext().ospHipRenderBatch(n, _handles(fbs), _handles(renderers), _handles(cameras), world.h,
                                   1 if graph else 0, ms)
    if rc != 0:
        check()
        raise OSPRayError("ospHipRenderBatch failed")
    return [float(x) for x in ms]


def tune_batch(fbs, renderers, cameras, world: Handle, reps: int = 2):
    """Run the batch autotuner; returns (best batch ms, (traceBlocksPerCU,
    shadeBlocksPerCU, pathBudgetMB, batchStreams, tileWidth))."""
    best = (C.c_int * 5)()
    n = len(fbs)
    ms = K.ext().ospHipTuneBatch(n, _handles(fbs), _handles(renderers), _handles(cameras), world.h, reps, best)
    if ms < 0:
        check()
        raise OSPRayError("ospHipTuneBatch failed")
    return float(ms), tuple(int(x) for x in best)


def volume_stats() -> tuple:
    """(voxel samples, macrocell intervals, tracked segments) of the
    free-path sampling, then of the transmittance estimates, of the last
    frame rendered with "traversalStats"."""
    out = (C.c_uint64 * 6)()
    K.ext().ospHipGetVolumeStats(out)
    return tuple(int(x) for x in out)


def trace_rays(world: Handle, rays8: np.ndarray):
    """Closest hits of rays (N x 8: org, dir, t0, t1) on the GPU.
    Returns (tuv N x 3, ids N x 3 [instance, model, primitive], (nodes, prims))."""
    r = np.ascontiguousarray(rays8, dtype=np.float32)
    n = r.shape[0]
    tuv = np.empty((n, 3), np.float32)
    ids = np.empty((n, 3), np.uint32)
    st = (C.c_uint64 * 2)()
    rc = K.ext().ospHipTraceRays(world.h, r.ctypes.data, n, tuv.ctypes.data, ids.ctypes.data, st)
    if rc < 0:
        check()
        raise OSPRayError("ospHipTraceRays failed")
    return tuv, ids, (int(st[0]), int(st[1]))


def device_info() -> tuple:
    name = C.create_string_buffer(64)
    cus = C.c_int(0)
    K.ext().ospHipGetDeviceInfo(name, 64, C.byref(cus))
    return name.value.decode(), cus.value


def world_info(world: Handle) -> tuple:
    out = (C.c_int64 * 4)()
    K.ext().ospHipGetWorldInfo(world.h, out)
    return tuple(int(x) for x in out)


def world_build(world: Handle) -> tuple:
    """(BLAS build milliseconds of the last commit, BLASes built on the
    device, traversal stack entries the BVH needs)."""
    ms, ndev, need = C.c_double(0.0), C.c_int(0), C.c_int(0)
    K.ext().ospHipGetWorldBuild(world.h, C.byref(ms), C.byref(ndev), C.byref(need))
    return float(ms.value), int(ndev.value), int(need.value)


def copy_bandwidth(nbytes: int, reps: int = 10) -> float:
    """Mean ms of one device-to-device copy of nbytes by the plain copy kernel."""
    ms = C.c_float(0.0)
    if K.ext().ospHipCopyBandwidth(nbytes, reps, C.byref(ms)) != 0:
        check()
        raise OSPRayError("ospHipCopyBandwidth failed")
    return float(ms.value)
