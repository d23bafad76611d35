"""ctypes bindings of the OSPRay C API served by ospray_amd/lib/libospray.so.

The functions keep the C names and argument meaning of
ospray/include/ospray/ospray.h and ospray_util.h; `ext` holds the hip
device's extensions (include/ospray/ospray_hip.h).  Loading fails loudly if
the native libraries have not been built (`python -m ospray_amd.build`).
"""
from __future__ import annotations

import ctypes as C
import os

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

# OSPDataType values (OSPEnums.h)
OSP_DEVICE, OSP_VOID_PTR, OSP_BOOL = 100, 200, 250
OSP_OBJECT = 0x8000000
(OSP_DATA, OSP_CAMERA, OSP_FRAMEBUFFER, OSP_FUTURE, OSP_GEOMETRIC_MODEL, OSP_GEOMETRY, OSP_GROUP,
 OSP_IMAGE_OPERATION, OSP_INSTANCE, OSP_LIGHT, OSP_MATERIAL, OSP_RENDERER, OSP_TEXTURE,
 OSP_TRANSFER_FUNCTION, OSP_VOLUME, OSP_VOLUMETRIC_MODEL, OSP_WORLD) = range(OSP_OBJECT + 100, OSP_OBJECT + 117)
OSP_STRING = 1500
OSP_UCHAR = 2500
OSP_INT, OSP_VEC2I, OSP_VEC3I, OSP_VEC4I = 4000, 4001, 4002, 4003
OSP_UINT, OSP_VEC2UI, OSP_VEC3UI, OSP_VEC4UI = 4500, 4501, 4502, 4503
OSP_FLOAT, OSP_VEC2F, OSP_VEC3F, OSP_VEC4F = 6000, 6001, 6002, 6003
OSP_BOX1F = 10000
OSP_AFFINE3F = 12003

# framebuffer formats / channels / events (OSPEnums.h)
OSP_FB_NONE, OSP_FB_RGBA8, OSP_FB_SRGBA, OSP_FB_RGBA32F = 0, 1, 2, 3
OSP_FB_COLOR, OSP_FB_DEPTH, OSP_FB_ACCUM, OSP_FB_VARIANCE = 1, 2, 4, 8
OSP_FB_NORMAL, OSP_FB_ALBEDO, OSP_FB_ID_PRIMITIVE, OSP_FB_ID_OBJECT = 16, 32, 64, 128
OSP_FB_ID_INSTANCE, OSP_FB_FIRST_NORMAL, OSP_FB_POSITION = 256, 512, 1024
OSP_NONE_FINISHED, OSP_WORLD_RENDERED, OSP_WORLD_COMMITTED, OSP_FRAME_FINISHED, OSP_TASK_FINISHED = (
    0, 10, 20, 30, 100000)
OSP_NO_ERROR, OSP_UNKNOWN_ERROR, OSP_INVALID_ARGUMENT, OSP_INVALID_OPERATION = 0, 1, 2, 3
OSP_OUT_OF_MEMORY, OSP_UNSUPPORTED_CPU, OSP_VERSION_MISMATCH = 4, 5, 6
OSP_PIXELFILTER_POINT, OSP_PIXELFILTER_BOX, OSP_PIXELFILTER_GAUSS = 0, 1, 2
OSP_INTENSITY_QUANTITY_RADIANCE, OSP_INTENSITY_QUANTITY_IRRADIANCE = 0, 1


class OSPBounds(C.Structure):
    _fields_ = [("lower", C.c_float * 3), ("upper", C.c_float * 3)]


class OSPPickResult(C.Structure):
    _fields_ = [("hasHit", C.c_int), ("worldPosition", C.c_float * 3), ("instance", C.c_void_p),
                ("model", C.c_void_p), ("primID", C.c_uint32)]


class OSPHipKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("totalMs", C.c_double)]


_VP, _CP, _U32, _U64, _I64, _F = C.c_void_p, C.c_char_p, C.c_uint32, C.c_uint64, C.c_int64, C.c_float

ERROR_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_char_p)
STATUS_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_char_p)

_SIGS = {
    "ospInit": (C.c_int, [C.POINTER(C.c_int), C.POINTER(_CP)]),
    "ospDeviceGetProperty": (_I64, [_VP, C.c_int]),
    "ospShutdown": (None, []),
    "ospNewDevice": (_VP, [_CP]),
    "ospSetCurrentDevice": (None, [_VP]),
    "ospGetCurrentDevice": (_VP, []),
    "ospDeviceSetParam": (None, [_VP, _CP, C.c_int, _VP]),
    "ospDeviceRemoveParam": (None, [_VP, _CP]),
    "ospDeviceSetErrorCallback": (None, [_VP, ERROR_CALLBACK, _VP]),
    "ospDeviceSetStatusCallback": (None, [_VP, STATUS_CALLBACK, _VP]),
    "ospDeviceGetLastErrorCode": (C.c_int, [_VP]),
    "ospDeviceGetLastErrorMsg": (_CP, [_VP]),
    "ospDeviceCommit": (None, [_VP]),
    "ospDeviceRelease": (None, [_VP]),
    "ospDeviceRetain": (None, [_VP]),
    "ospLoadModule": (C.c_int, [_CP]),
    "ospNewSharedData": (_VP, [_VP, C.c_int, _U64, _I64, _U64, _I64, _U64, _I64, _VP, _VP]),
    "ospNewData": (_VP, [C.c_int, _U64, _U64, _U64]),
    "ospCopyData": (None, [_VP, _VP, _U64, _U64, _U64]),
    "ospNewLight": (_VP, [_CP]),
    "ospNewCamera": (_VP, [_CP]),
    "ospNewGeometry": (_VP, [_CP]),
    "ospNewVolume": (_VP, [_CP]),
    "ospNewGeometricModel": (_VP, [_VP]),
    "ospNewVolumetricModel": (_VP, [_VP]),
    "ospNewMaterial": (_VP, [_CP]),
    "ospNewTransferFunction": (_VP, [_CP]),
    "ospNewTexture": (_VP, [_CP]),
    "ospNewGroup": (_VP, []),
    "ospNewInstance": (_VP, [_VP]),
    "ospNewWorld": (_VP, []),
    "ospGetBounds": (OSPBounds, [_VP]),
    "ospSetParam": (None, [_VP, _CP, C.c_int, _VP]),
    "ospRemoveParam": (None, [_VP, _CP]),
    "ospCommit": (None, [_VP]),
    "ospRelease": (None, [_VP]),
    "ospRetain": (None, [_VP]),
    "ospNewFrameBuffer": (_VP, [C.c_int, C.c_int, C.c_int, _U32]),
    "ospNewImageOperation": (_VP, [_CP]),
    "ospMapFrameBuffer": (_VP, [_VP, C.c_int]),
    "ospUnmapFrameBuffer": (None, [_VP, _VP]),
    "ospGetVariance": (_F, [_VP]),
    "ospResetAccumulation": (None, [_VP]),
    "ospNewRenderer": (_VP, [_CP]),
    "ospRenderFrame": (_VP, [_VP, _VP, _VP, _VP]),
    "ospIsReady": (C.c_int, [_VP, C.c_int]),
    "ospWait": (None, [_VP, C.c_int]),
    "ospCancel": (None, [_VP]),
    "ospGetProgress": (_F, [_VP]),
    "ospGetTaskDuration": (_F, [_VP]),
    "ospPick": (None, [C.POINTER(OSPPickResult), _VP, _VP, _VP, _VP, _F, _F]),
    # ospray_util.h
    "ospNewSharedData1D": (_VP, [_VP, C.c_int, _U64]),
    "ospNewData1D": (_VP, [C.c_int, _U64]),
    "ospCopyData1D": (None, [_VP, _VP, _U64]),
    "ospSetString": (None, [_VP, _CP, _CP]),
    "ospSetObject": (None, [_VP, _CP, _VP]),
    "ospSetBool": (None, [_VP, _CP, C.c_int]),
    "ospSetFloat": (None, [_VP, _CP, _F]),
    "ospSetInt": (None, [_VP, _CP, C.c_int]),
    "ospSetUInt": (None, [_VP, _CP, C.c_uint]),
    "ospSetVec2f": (None, [_VP, _CP, _F, _F]),
    "ospSetVec3f": (None, [_VP, _CP, _F, _F, _F]),
    "ospSetVec4f": (None, [_VP, _CP, _F, _F, _F, _F]),
    "ospSetObjectAsData": (None, [_VP, _CP, C.c_int, _VP]),
    "ospRenderFrameBlocking": (_F, [_VP, _VP, _VP, _VP]),
}

_EXT_SIGS = {
    "ospHipGetKernelStats": (C.c_int, [C.POINTER(OSPHipKernelStat), C.c_int]),
    "ospHipResetKernelStats": (C.c_int, []),
    "ospHipGetTraversalStats": (C.c_int, [C.POINTER(_U64)]),
    "ospHipGetVolumeStats": (C.c_int, [C.POINTER(_U64)]),
    "ospHipRenderBatch": (C.c_int, [C.c_int, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP), _VP, C.c_int,
                                    C.POINTER(C.c_float)]),
    "ospHipTuneBatch": (C.c_float, [C.c_int, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP), _VP, C.c_int,
                                    C.POINTER(C.c_int)]),
    "ospHipGetNrcState": (C.c_int, [_VP, C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(_I64),
                                     C.POINTER(_I64)]),
    "ospHipTraceRays": (C.c_int, [_VP, _VP, C.c_int, _VP, _VP, C.POINTER(_U64)]),
    "ospHipGetDeviceInfo": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "ospHipGetWorldInfo": (C.c_int, [_VP, C.POINTER(_I64)]),
    "ospHipGetWorldBuild": (C.c_int, [_VP, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ospHipCopyBandwidth": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(_F)]),
}


class _Lib:
    def __init__(self, path, sigs):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build the native libraries first "
                               "(python -m ospray_amd.build)")
        self.path = path
        self.dll = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, (res, args) in sigs.items():
            fn = getattr(self.dll, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)


_lib = None
_ext = None


def lib() -> _Lib:
    """The C API (libospray.so)."""
    global _lib
    if _lib is None:
        _lib = _Lib(os.path.join(LIBDIR, "libospray.so"), _SIGS)
    return _lib


def ext() -> _Lib:
    """Extensions of the hip device (libospray_module_hip.so, loaded by
    ospLoadModule("hip"))."""
    global _ext
    if _ext is None:
        _ext = _Lib(os.path.join(LIBDIR, "libospray_module_hip.so"), _EXT_SIGS)
    return _ext
