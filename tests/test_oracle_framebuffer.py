"""Frame-buffer semantics of the oracle that no image pins: adaptive
accumulation (fb/LocalFB.cpp:238-261, fb/FrameBuffer.h:183-191).

From frameID >= max(2, 16/spp) on, with a variance buffer, a frame renders
only the 8x8 tasks whose last error (Variance.ispc:11-59, computed on odd
frames) exceeds `varianceThreshold`; the other tasks keep their buffers.
"""
import numpy as np

from ospray_amd import scenes
from oracle import orc

CH = 1 | 4 | 8  # colour, accum, variance


def _frames(sc, rp, n):
    osc = orc.OracleScene(sc)
    fb = orc.OracleFrameBuffer(sc.width, sc.height, 3, CH)
    p = dict(sc.renderer)
    p.update(rp)
    cam, ren = orc.make_camera(sc.camera), orc.make_renderer(p)
    out = []
    for _ in range(n):
        fb.render(osc, cam, ren, 8)
        out.append(fb.channel(1).copy())
    return out, fb, ren


def test_huge_threshold_freezes_after_minimum_frames():
    sc = scenes.c0(40, 32, spp=4)  # minimum adaptive frames: max(2, 16/4) = 4
    imgs, _, _ = _frames(sc, {"varianceThreshold": 1e30}, 7)
    assert not np.array_equal(imgs[2], imgs[3])  # frame 3 still renders every task
    for k in (4, 5, 6):
        assert np.array_equal(imgs[k], imgs[3])


def test_tiny_threshold_equals_plain_accumulation():
    sc = scenes.c0(40, 32, spp=4)
    a, _, _ = _frames(sc, {"varianceThreshold": 1e-30}, 6)
    b, _, _ = _frames(sc, {}, 6)
    assert np.array_equal(a[-1], b[-1])


def test_only_tasks_above_threshold_change():
    sc = scenes.c0(48, 48, spp=8)  # minimum adaptive frames: 2
    # the task errors after frame 1 (odd: variance computed); the threshold
    # between them so that both kinds of task exist
    _, fb, _ = _frames(sc, {}, 2)
    errs = fb.task_variance()
    assert errs.size == 36
    thr = float(np.median(errs))
    imgs, _, _ = _frames(sc, {"varianceThreshold": thr}, 3)
    before, after = imgs[1], imgs[2]
    changed = np.zeros((6, 6), bool)
    for ty in range(6):
        for tx in range(6):
            changed[ty, tx] = not np.array_equal(before[ty * 8:ty * 8 + 8, tx * 8:tx * 8 + 8],
                                                 after[ty * 8:ty * 8 + 8, tx * 8:tx * 8 + 8])
    active = errs.reshape(6, 6) > np.float32(thr)
    assert active.any() and (~active).any()
    assert np.array_equal(changed, active)
