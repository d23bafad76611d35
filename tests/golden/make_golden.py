"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

  c0_hits.npz       1,000 camera rays of the C0 scene (32x32 camera) with
                    the oracle's closest hits (t, u, v; instance, model,
                    primitive) and the FP64 brute-force ids and t
  c0_frames.npz     the C0 scene at 32x32 after 2 accumulated frames: every
                    frame-buffer channel and the variance estimate
  c1small_frames.npz the reduced 1080p scene (48x27, 2 spp, 97^2 grid, 40
                    instances: Lambert, metal, glass, sun cone, quad lights)
                    after 2 frames: colour, depth and the three id channels
  ld_samples.npz    LDSampler 5-D samples (pixel sample, lens, time) for 8
                    seeds x 64 sample indices

The fixtures are data: inputs and the oracle's outputs.  The CPU tests
check that the oracle still reproduces them (and what closed forms pin
about them); the GPU tests check the device against them without running
the oracle.  Run: python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ospray_amd import scenes  # noqa: E402
from oracle import orc  # noqa: E402

ALL_CH = 1 | 2 | 4 | 8 | 16 | 32 | 64 | 128 | 256 | 512 | 1024
FRAME_CH = {"color": 1, "depth": 2, "normal": 16, "albedo": 32, "prim": 64, "obj": 128, "inst": 256,
            "first_normal": 512, "position": 1024}


def camera_rays(sc, n, seed):
    """Rays through uniformly random points of the camera's image plane."""
    rng = np.random.default_rng(seed)
    cam = sc.camera
    pos = np.array(cam["position"], np.float64)
    d = np.array(cam["direction"], np.float64)
    d /= np.linalg.norm(d)
    du = np.cross(d, np.array(cam["up"], np.float64))
    du /= np.linalg.norm(du)
    dv = np.cross(du, d)
    h = np.tan(np.radians(cam["fovy"]) / 2)
    a = cam.get("aspect", 1.0)
    sx, sy = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    dirs = d + (sx * h * a)[:, None] * du + (sy * h)[:, None] * dv
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return np.concatenate([np.tile(pos, (n, 1)), dirs, np.zeros((n, 1)), np.full((n, 1), np.inf)],
                          1).astype(np.float32)


def frames(sc, nframes, channels):
    fb, _, oren = orc.render_scene(sc, frames=nframes, fmt=3, channels=channels)
    out = {k: fb.channel(ch) for k, ch in FRAME_CH.items() if channels & ch}
    if channels & 8:
        out["variance_estimate"] = np.float32(fb.variance(oren))
    return out


def main():
    sc = scenes.c0(32, 32)
    osc = orc.OracleScene(sc)
    rays = camera_rays(sc, 1000, 7)
    tuv, ids, _ = osc.intersect(rays)
    ftuv, fids, _ = osc.intersect(rays, fp64=True)
    np.savez_compressed(os.path.join(HERE, "c0_hits.npz"), rays=rays, tuv=tuv, ids=ids, fp64_t=ftuv[:, 0],
                        fp64_ids=fids)
    np.savez_compressed(os.path.join(HERE, "c0_frames.npz"), **frames(sc, 2, ALL_CH))
    c1 = scenes.c1(width=48, height=27, spp=2, grid=97, instances=40)
    np.savez_compressed(os.path.join(HERE, "c1small_frames.npz"), **frames(c1, 2, 1 | 2 | 4 | 64 | 128 | 256))
    seeds = np.arange(8, dtype=np.uint32) * np.uint32(0x9E3779B9)
    ld = np.empty((8, 64, 5), np.float32)
    buf = (C.c_float * 5)()
    for a, seed in enumerate(seeds):
        for i in range(64):
            orc.lib().orc_ld_samples5(int(seed), i, 16, buf)
            ld[a, i] = np.frombuffer(buf, np.float32)
    np.savez_compressed(os.path.join(HERE, "ld_samples.npz"), seeds=seeds, samples=ld)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
