"""hpt_trace_rays_noalloc (include/ospray_hip/hip_pt.h), the closest-hit
query whose traversal scratch rows are a fixed device array of the module,
against hpt_trace_rays with a caller-allocated scratch, both through the C
ABI on the same committed world and device-resident rays: hits (t, u, v and
ids) and node/primitive counts equal bit for bit on two calls in a row, the
second with fewer rays and another resident grid (the array keeps the first
call's contents); sizes beyond the array are refused before any launch.
The HIP runtime calls here (allocation, copies) are the test's own: the
entry point under test allocates nothing."""
import ctypes as C
import os

import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODULE = os.path.join(ROOT, "ospray_amd", "lib", "libospray_module_hip.so")
H2D, D2H = 1, 2
HIP_ERROR_INVALID_VALUE = 1


def _libs():
    mod = C.CDLL(MODULE)
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the module is linked against
    vp = C.c_void_p
    mod.ospHipGetWorldLaunchScene.restype = vp
    mod.ospHipGetWorldLaunchScene.argtypes = [vp]
    mod.hpt_trace_rays.restype = C.c_int
    mod.hpt_trace_rays.argtypes = [vp, vp, C.c_int, vp, vp, vp, C.c_int, vp, vp]
    mod.hpt_trace_rays_noalloc.restype = C.c_int
    mod.hpt_trace_rays_noalloc.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int, vp, vp]
    for f in (mod.hpt_trace_noalloc_max_blocks, mod.hpt_trace_block_size, mod.hpt_overflow_depth):
        f.restype = C.c_int
        f.argtypes = []
    hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
    for f in (hip.hipMalloc, hip.hipFree, hip.hipMemcpy, hip.hipMemset, hip.hipDeviceSynchronize):
        f.restype = C.c_int
    return mod, hip


class _Buf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(nbytes, 16)) == 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _rays(sc, n, seed):
    rng = np.random.default_rng(seed)
    cam = sc.camera
    pos = np.array(cam["position"], np.float64)
    d = np.array(cam["direction"], np.float64)
    d /= np.linalg.norm(d)
    du = np.cross(d, np.array(cam["up"], np.float64))
    du /= np.linalg.norm(du)
    dv = np.cross(du, d)
    h = np.tan(np.radians(cam["fovy"]) / 2)
    sx, sy = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    dirs = d + (sx * h)[:, None] * du + (sy * h)[:, None] * dv
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return np.concatenate([np.tile(pos, (n, 1)), dirs, np.zeros((n, 1)), np.full((n, 1), np.inf)],
                          1).astype(np.float32)


def test_noalloc_trace_equals_trace_rays(hip_device):
    mod, hip = _libs()
    sc = scenes.c0(32, 32)
    s = scenes.to_ospray(sc)
    scene = mod.ospHipGetWorldLaunchScene(s.world.h)
    assert scene
    maxb = mod.hpt_trace_noalloc_max_blocks()
    assert maxb >= 1
    block, depth = mod.hpt_trace_block_size(), mod.hpt_overflow_depth()
    ref_blocks = 2048
    ovf = _Buf(hip, ref_blocks * block * depth * 4)
    assert hip.hipMemset(ovf.p, 0, ref_blocks * block * depth * 4) == 0
    for n, blocks, seed in ((60000, maxb, 5), (7001, max(1, maxb // 3), 6)):
        print(f"noalloc: n={n} residentBlocks={blocks}", flush=True)
        rays = _rays(sc, n, seed)
        d_rays = _Buf(hip, rays.nbytes)
        d_rays.put(rays)
        outs = [(_Buf(hip, n * 12), _Buf(hip, n * 12), _Buf(hip, 16)) for _ in range(2)]
        for _, _, st in outs:
            st.put(np.zeros(2, np.uint64))
        (ta, ia, sa), (tb, ib, sb) = outs
        assert mod.hpt_trace_rays(scene, d_rays.p, n, ta.p, ia.p, ovf.p, ref_blocks, sa.p, None) == 0
        assert mod.hpt_trace_rays_noalloc(scene, d_rays.p, n, tb.p, ib.p, blocks, sb.p, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        tuv_a, tuv_b = ta.get((n, 3), np.uint32), tb.get((n, 3), np.uint32)
        ids_a, ids_b = ia.get((n, 3), np.uint32), ib.get((n, 3), np.uint32)
        assert np.array_equal(tuv_a, tuv_b)
        assert np.array_equal(ids_a, ids_b)
        assert np.any(ids_a[:, 0] != 0xFFFFFFFF) and np.any(ids_a[:, 0] == 0xFFFFFFFF)
        st_a, st_b = sa.get(2, np.uint64), sb.get(2, np.uint64)
        assert np.array_equal(st_a, st_b) and st_a[0] > 0
        # the same hits as the device's allocating query
        tuv_ref, ids_ref, _ = A.trace_rays(s.world, rays)
        assert np.array_equal(ids_b, ids_ref.astype(np.uint32))
        assert np.array_equal(tuv_b.view(np.float32), tuv_ref.astype(np.float32))
        for b in (d_rays, ta, ia, sa, tb, ib, sb):
            b.free()
    ovf.free()
    # sizes beyond the array, or no grid, are refused before launching
    assert mod.hpt_trace_rays_noalloc(scene, None, 0, None, None, maxb + 1, None, None) == HIP_ERROR_INVALID_VALUE
    assert mod.hpt_trace_rays_noalloc(scene, None, 0, None, None, 0, None, None) == HIP_ERROR_INVALID_VALUE
    assert hip.hipDeviceSynchronize() == 0
