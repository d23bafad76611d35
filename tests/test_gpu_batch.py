"""Batched views (ospHipRenderBatch, ospHipTuneBatch): a batch of views of
one world rendered in one enqueue, eagerly or as one captured HIP graph, on
concurrent lanes or on one stream, gives each view exactly the image
ospRenderFrame gives it; every call renders one frame of each view, so
relaunches of a captured batch accumulate; the autotuner's settings leave
the images unchanged."""
import numpy as np
import pytest

from ospray_amd import api as A
from ospray_amd import capi as K
from ospray_amd import scenes

pytestmark = pytest.mark.gpu


def _views(s, specs):
    fbs, rens, cams = [], [], []
    for (w, h, spp, pos) in specs:
        cams.append(A.new("camera", "perspective", {"position": pos, "direction": tuple(-x for x in pos),
                                                     "up": (0.0, 1.0, 0.0), "fovy": 45.0, "aspect": w / h}))
        rens.append(A.new("renderer", "pathtracer", {"pixelSamples": spp, "maxPathLength": 4}))
        fbs.append(A.framebuffer(w, h, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_DEPTH))
    return fbs, rens, cams


SPECS = [(40, 24, 1, (0.0, 4.0, -20.0)), (64, 64, 2, (15.0, 6.0, -10.0)), (24, 40, 4, (-12.0, 10.0, 12.0))]


def _set_batch_streams(n):
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "batchStreams", n)
    L.ospDeviceCommit(dev)


def _images(fbs):
    return [A.map_color(fb, w, h, K.OSP_FB_RGBA32F) for fb, (w, h, _, _) in zip(fbs, SPECS)]


def test_batch_equals_single_frames(hip_device):
    sc = scenes.c0(32, 32)
    s = scenes.to_ospray(sc)
    fbs, rens, cams = _views(s, SPECS)
    single = []
    for fb, r, c, (w, h, _, _) in zip(fbs, rens, cams, SPECS):
        A.render(fb, r, c, s.world)
        single.append(A.map_color(fb, w, h, K.OSP_FB_RGBA32F))
    # the views on concurrent lanes (the default), then on one stream
    ms = A.render_batch(fbs, rens, cams, s.world, graph=False)
    assert all(a > 0 for a in ms)
    for a, b in zip(_images(fbs), single):
        assert np.array_equal(a, b)
    for _ in range(3):  # capture, then two relaunches
        A.render_batch(fbs, rens, cams, s.world, graph=True)
        for a, b in zip(_images(fbs), single):
            assert np.array_equal(a, b)
    _set_batch_streams(1)
    try:
        ms = A.render_batch(fbs, rens, cams, s.world, graph=False)
        assert all(a > 0 for a in ms) and ms == sorted(ms)  # views finish in order on one stream
        for a, b in zip(_images(fbs), single):
            assert np.array_equal(a, b)
        A.render_batch(fbs, rens, cams, s.world, graph=True)
        for a, b in zip(_images(fbs), single):
            assert np.array_equal(a, b)
    finally:
        _set_batch_streams(4)
    # a recommitted camera invalidates the captured graph
    A.set_param(cams[1].h, "position", (16.0, 6.0, -10.0))
    A.commit(cams[1])
    A.render_batch(fbs, rens, cams, s.world, graph=True)
    moved = _images(fbs)[1]
    fb = A.framebuffer(64, 64, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_DEPTH)
    A.render(fb, rens[1], cams[1], s.world)
    assert np.array_equal(moved, A.map_color(fb, 64, 64, K.OSP_FB_RGBA32F))
    assert not np.array_equal(moved, single[1])


def test_autotuner_keeps_images(hip_device):
    sc = scenes.c0(32, 32)
    s = scenes.to_ospray(sc)
    fbs, rens, cams = _views(s, SPECS)
    A.render_batch(fbs, rens, cams, s.world, graph=False)
    before = _images(fbs)
    ms, best = A.tune_batch(fbs, rens, cams, s.world, reps=1)
    assert ms > 0 and best[0] in (8, 12, 16) and best[1] in (4, 8, 16) and best[2] in (2048, 16384)
    assert best[3] in (1, 2, 4, 8) and best[4] in (4, 8, 16, 32)
    A.render_batch(fbs, rens, cams, s.world, graph=True)
    for a, b in zip(_images(fbs), before):
        assert np.array_equal(a, b)
    # back to the defaults for the tests that follow
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    for k, v in (("traceBlocksPerCU", 12), ("shadeBlocksPerCU", 8), ("pathBudgetMB", 16384), ("batchStreams", 4),
                 ("tileWidth", 8)):
        A.set_device_param(dev, k, v)
    L.ospDeviceCommit(dev)


def test_graph_relaunches_accumulate(hip_device):
    """Accumulating framebuffers (with a variance buffer, one renderer with
    adaptive accumulation): every call renders one frame of every view (the
    first eagerly, while it captures the graph; each later one by a
    relaunch), as ospRenderFrame would; images, depth and ids and
    ospGetVariance (which follows the variance buffer and, with the
    threshold, the adaptive task choice) equal those of as many eager
    frames."""
    sc = scenes.c0(32, 32)
    s = scenes.to_ospray(sc)
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_VARIANCE | K.OSP_FB_DEPTH | K.OSP_FB_ID_OBJECT
    specs = [(40, 24, 1, (0.0, 4.0, -20.0), 0.0), (64, 64, 8, (15.0, 6.0, -10.0), 0.05),
             (24, 40, 4, (-12.0, 10.0, 12.0), 0.0)]

    def views():
        fbs, rens, cams = [], [], []
        for (w, h, spp, pos, thr) in specs:
            cams.append(A.new("camera", "perspective", {"position": pos, "direction": tuple(-x for x in pos),
                                                         "up": (0.0, 1.0, 0.0), "fovy": 45.0, "aspect": w / h}))
            rens.append(A.new("renderer", "pathtracer", {"pixelSamples": spp, "maxPathLength": 4,
                                                         "varianceThreshold": thr}))
            fbs.append(A.framebuffer(w, h, K.OSP_FB_RGBA32F, ch))
        return fbs, rens, cams

    fbs, rens, cams = views()
    efbs, erens, ecams = views()
    L = K.lib()
    frames = 0
    for call in range(4):
        A.render_batch(fbs, rens, cams, s.world, graph=True)
        for fb, r, c in zip(efbs, erens, ecams):
            A.render(fb, r, c, s.world)
        frames += 1
        for fb, efb, (w, h, *_rest) in zip(fbs, efbs, specs):
            assert np.array_equal(A.map_color(fb, w, h, K.OSP_FB_RGBA32F), A.map_color(efb, w, h, K.OSP_FB_RGBA32F))
            for c in (K.OSP_FB_DEPTH, K.OSP_FB_ID_OBJECT):
                assert np.array_equal(A.map_channel(fb, c, w, h), A.map_channel(efb, c, w, h)), (call, c)
            assert L.ospGetVariance(fb.h) == L.ospGetVariance(efb.h), (call, frames)
    assert frames == 4
    # the images do change from launch to launch (accumulation advances)
    before = A.map_color(fbs[0], 40, 24, K.OSP_FB_RGBA32F)
    A.render_batch(fbs, rens, cams, s.world, graph=True)
    assert not np.array_equal(before, A.map_color(fbs[0], 40, 24, K.OSP_FB_RGBA32F))


def test_c2_views_against_oracle_bands(hip_device):
    """The 64-view C2 batch of the full 1080p scene (seeded positions on a
    radius-30 sphere, resolutions {256^2, 512^2, 768x432, 1024^2}, spp
    {1, 2, 4, 8}, maxPathLength 8) on accumulating framebuffers, one frame
    rendered while the graph is captured and one by its launch: every view's two middle rows equal the
    oracle's render of those rows over the same two frames, bit for bit."""
    from oracle import orc

    sc = scenes.c1()
    views = scenes.c2_views()
    s = scenes.to_ospray(sc)
    fbs, rens, cams = [], [], []
    for v in views:
        cams.append(A.new("camera", "perspective", dict(v["camera"])))
        rens.append(A.new("renderer", "pathtracer", {"pixelSamples": v["spp"], "maxPathLength": 8}))
        fbs.append(A.framebuffer(v["width"], v["height"], K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_ACCUM))
    A.render_batch(fbs, rens, cams, s.world, graph=True)  # frame 0, eager, while capturing
    A.render_batch(fbs, rens, cams, s.world, graph=True)  # frame 1, the graph's first launch
    osc = orc.OracleScene(sc)
    for v, fb in zip(views, fbs):
        w, h = v["width"], v["height"]
        img = A.map_color(fb, w, h, K.OSP_FB_RGBA32F)
        assert np.all(np.isfinite(img))
        y0 = h // 2
        ofb = orc.OracleFrameBuffer(w, h, 3, K.OSP_FB_COLOR | K.OSP_FB_ACCUM)
        cam = orc.make_camera(v["camera"])
        ren = orc.make_renderer({"pixelSamples": v["spp"], "maxPathLength": 8})
        for _ in range(2):
            ofb.render(osc, cam, ren, 16, (y0, y0 + 2))
        assert np.array_equal(img[y0:y0 + 2], ofb.channel(K.OSP_FB_COLOR)[y0:y0 + 2]), v
