"""Futures of the hip device (RenderingFuture.h:60-103):

  * ospGetProgress rises with the frame's passes and is 1 once the frame
    has finished (FrameBuffer::getCurrentProgress, FrameBuffer.cpp:169-178);
  * ospCancel stops the frame's remaining passes and bounces
    (FrameBuffer::cancelFrame, FrameBuffer.cpp:180-187; the render tasks'
    check, PathTracer.ispc:231): the cancelled frame finishes early and
    leaves the framebuffer's image as it was, and the next frame renders;
  * cancelling a frame that has already finished changes nothing: later
    frames equal those of a framebuffer whose frames were never cancelled.

The long frame is C0 at 2048x2048 with 16 spp and a 1 MiB path budget, so
that every pass holds one sample per pixel (16 passes of 4 M paths: few
enough launches that the host finishes enqueuing long before the GPU
finishes the frame)."""
import time

import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes

pytestmark = pytest.mark.gpu

W = H = 2048
SPP = 16


def _budget(mb):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "pathBudgetMB", mb)
    L.ospDeviceCommit(dev)


@pytest.fixture(scope="module")
def long_frame(hip_device):
    sc = scenes.c0(W, H)
    s = scenes.to_ospray(sc, {"pixelSamples": SPP})
    _budget(1)
    yield s
    _budget(16384)


def _start(fb, s):
    L = K.lib()
    f = L.ospRenderFrame(fb.h, s.renderer.h, s.camera.h, s.world.h)
    if not f:
        A.check()
    assert f
    return f


def _finish(f):
    L = K.lib()
    L.ospWait(f, K.OSP_TASK_FINISHED)
    p = L.ospGetProgress(f)
    L.ospRelease(f)
    return p


def test_progress_rises_to_one(long_frame):
    L = K.lib()
    fb = A.framebuffer(W, H, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_ACCUM)
    f = _start(fb, long_frame)
    seen = []
    while not L.ospIsReady(f, K.OSP_TASK_FINISHED):
        seen.append(L.ospGetProgress(f))
        time.sleep(0.001)
    assert _finish(f) == 1.0
    assert seen == sorted(seen)
    assert all(0.0 <= p <= 1.0 for p in seen)
    assert any(0.0 < p < 1.0 for p in seen), seen[:8]


def test_cancel_stops_the_frame(long_frame):
    L = K.lib()
    fb = A.framebuffer(W, H, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_DEPTH)
    t = time.perf_counter()
    _finish(_start(fb, long_frame))
    full_s = time.perf_counter() - t
    before = A.map_color(fb, W, H, K.OSP_FB_RGBA32F)
    depth = A.map_channel(fb, K.OSP_FB_DEPTH, W, H)
    t = time.perf_counter()
    f = _start(fb, long_frame)
    L.ospCancel(f)
    assert _finish(f) == 1.0
    cancelled_s = time.perf_counter() - t
    print(f"full frame {full_s * 1e3:.1f} ms, cancelled frame {cancelled_s * 1e3:.1f} ms")
    assert cancelled_s < 0.5 * full_s
    assert np.array_equal(A.map_color(fb, W, H, K.OSP_FB_RGBA32F), before)
    assert np.array_equal(A.map_channel(fb, K.OSP_FB_DEPTH, W, H), depth)
    # the next frame is not cancelled: it accumulates
    _finish(_start(fb, long_frame))
    after = A.map_color(fb, W, H, K.OSP_FB_RGBA32F)
    assert np.all(np.isfinite(after))
    assert not np.array_equal(after, before)


def test_cancel_after_the_frame_changes_nothing(hip_device):
    L = K.lib()
    sc = scenes.c0(64, 64)
    s = scenes.to_ospray(sc, {"pixelSamples": 2})
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_VARIANCE
    a = A.framebuffer(64, 64, K.OSP_FB_RGBA32F, ch)
    b = A.framebuffer(64, 64, K.OSP_FB_RGBA32F, ch)
    for _ in range(3):
        f = _start(a, s)
        L.ospWait(f, K.OSP_TASK_FINISHED)
        L.ospCancel(f)
        L.ospRelease(f)
        _finish(_start(b, s))
    assert np.array_equal(A.map_color(a, 64, 64, K.OSP_FB_RGBA32F), A.map_color(b, 64, 64, K.OSP_FB_RGBA32F))
    assert L.ospGetVariance(a.h) == L.ospGetVariance(b.h)
