#!/usr/bin/env python3
"""Path samples/s of the hip `pathtracer` at 1080p.

Workload: the 4.2M-triangle scene at 1920x1080, 4 spp, maxPathLength 8
(seeded fBm heightfield 1448^2 + 256 icosphere instances, 60/25/15 %
diffuse/metal/glass, a sun and 4 quad lights; ospray_amd/scenes.py c1()).
A step is one ospRenderFrame + ospWait of that frame through the OSPRay C
API, the scene already committed and resident in HBM.  Prints one JSON line.

  value       whole-job path samples/s (1920*1080*4 per step / wall time)
  roofline    the dominant kernel (closest-hit traversal of the extension or
              the shadow rays) against MI355X's 8 TB/s: algorithmic bytes per
              launch = 64 B per BVH node visit + 48 B per primitive test +
              56 B per ray (32 B ray in, 24 B hit out), with node and
              primitive counts from the device's traversal counters on one
              extra frame, over the kernel's mean launch duration from HIP
              events recorded on the device stream during the timed steps
              (frac_alg); the measured DRAM bytes per launch (FETCH_SIZE +
              WRITE_SIZE from the committed rocprofv3 summary) over the same
              duration (frac_dram); frac = the lower of the two; `valu` =
              the kernel's VALU issue rate and active-lane fraction from the
              committed SQ-counter summary
  cpu_baseline the CPU oracle (a C++ restatement of the reference path
              tracer, built here with -O3 -march=native) on a bounded sample
              of the same frame's rows, on every CPU of the job's share, with
              the host's nproc, model name and a probe for an installed
              OSPRay/Embree
  c3          plain vs radiance-cache time to equal RMSE on the same scene:
              1 spp frames x 256, the cache after bounce 2, 2^16 training
              paths per frame, against a 4096-spp plain render (--no-c3
              skips it)

--workload c4 measures the mixed surface + volume scene instead (the C0
spheres inside a 512^3 fBm density volume, 1280x720, 2 spp, maxPathLength
10; scenes.c4()); its roofline line also covers the free-path kernel
k_volume: 32 B per voxel sample (the trilinear footprint) + 8 B per
macrocell interval + 80 B per queued ray (ray, hit record, sampler state).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path samples/s at 1080p; plain vs neural-cache time-to-equal-RMSE"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one traversal: 64 B per node visit and 48 B per
# primitive test (the declared record sizes of the metric's definition),
# plus 56 B per ray (32 B ray in, 24 B hit out).  The device's 4-wide node
# is one 128-byte line; NODE_RECORD_BYTES gives that figure beside it.
NODE_BYTES, PRIM_BYTES, RAY_BYTES = 64, 48, 56
NODE_RECORD_BYTES = 128
# volume tracking: 32 B per voxel sample (8 corner densities of the
# trilinear stencil), 8 B per macrocell interval (its majorant), 80 B per
# extension ray k_volume takes (ray in, hit in, path state out) and 32 B per
# transmittance estimate of a shadow record in k_resolve (its ray)
SAMPLE_BYTES, CELL_BYTES, VOL_RAY_BYTES, TRACK_BYTES = 32, 8, 80, 32
TRAFFIC_FILES = {"c1": os.path.join(ROOT, "profiles", "hbm_traffic_pathtracer_1080p.md"),
                 "c4": os.path.join(ROOT, "profiles", "hbm_traffic_pathtracer_c4_volume.md")}
VALU_FILE = os.path.join(ROOT, "profiles", "valu_traversal_pathtracer_1080p.md")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpus() -> dict:
    """The host's CPUs: nproc, the CPUs this process may run on, the CPU
    share the job is given (OMP_NUM_THREADS when set, which the one-GPU job
    environment sets to its CPU allotment) and the lscpu model name."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    share = allowed
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count() or 1, "allowed": allowed, "threads": max(1, share), "model": model}


def reference_probe() -> dict:
    """Looks for an installed OSPRay 3.x / Embree 4 / rkcommon / OpenVKL on
    this host (the reference CPU module would be timed against it)."""
    import glob
    import subprocess

    found = []
    try:
        r = subprocess.run(["ldconfig", "-p"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                           timeout=20)
        found += [l.strip() for l in r.stdout.splitlines() if re.search(r"ospray|embree4|rkcommon|openvkl", l)]
    except (OSError, subprocess.SubprocessError):
        pass
    for pat in ("/usr/lib*/libospray_module_cpu*", "/opt/*/lib*/libospray*", "/usr/local/lib*/libospray*"):
        found += glob.glob(pat)
    return {"ldconfig_and_paths": found, "reference_installed": bool(found)}


def build_native_oracle() -> str:
    """Builds the oracle with -O3 -march=native for this host (the timing
    build; same arithmetic as the test build) and returns its path."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"], check=True)
    return os.path.join(ROOT, "oracle", "liborc_native.so")


def cpu_baseline(sc, target_s: float) -> dict:
    """Oracle samples/s on evenly spaced 8-row bands of the frame (about
    target_s seconds of CPU work), on every CPU of the job's share, from the
    -O3 -march=native build."""
    t = time.perf_counter()
    os.environ["ORC_LIB"] = build_native_oracle()
    log(f"oracle built for this host (-O3 -march=native) in {time.perf_counter() - t:.1f} s")
    from oracle import orc

    cpus = host_cpus()
    probe = reference_probe()
    log(f"host CPUs: {cpus}; installed reference: {probe}")
    t = time.perf_counter()
    osc = orc.OracleScene(sc)
    log(f"oracle BVH built in {time.perf_counter() - t:.1f} s")
    cam, ren = orc.make_camera(sc.camera), orc.make_renderer(sc.renderer)
    fb = orc.OracleFrameBuffer(sc.width, sc.height, 3, 1 | 4)
    nt = cpus["threads"]
    nbands = (sc.height + 7) // 8
    spp = sc.renderer["pixelSamples"]

    def band(b):
        y0 = b * 8
        y1 = min(sc.height, y0 + 8)
        t0 = time.perf_counter()
        fb.render(osc, cam, ren, nt, (y0, y1))
        return time.perf_counter() - t0, (y1 - y0) * sc.width * spp

    # calibrate on the middle band, then spread the sample over the frame;
    # whole frames are repeated (successive accumulation frames) while the
    # sample is shorter than target_s
    dt, _ = band(nbands // 2)
    count = int(max(2, target_s / max(dt, 1e-4)))
    frames, count = divmod(count, nbands) if count >= nbands else (0, count)
    picks = list(range(nbands)) * frames
    if count >= 2:
        picks += sorted({int(round(i * (nbands - 1) / (count - 1))) for i in range(count)})
    total_t = total_n = 0
    for b in picks:
        dt, n = band(b)
        total_t += dt
        total_n += n
    log(f"oracle: {total_n} samples in {total_t:.1f} s on {nt} threads")
    return {"value": round(total_n / total_t, 1), "unit": "samples/s", "cores": nt, "kind": "port",
            "sample": f"{len(picks)} 8-row bands ({frames} whole frames + {len(picks) - frames * nbands} bands spread "
                      f"over the frame) of the {sc.width}x{sc.height} {spp}-spp frame: {total_n} path samples in "
                      f"{total_t:.1f} s; "
                      f"oracle C++ restatement of the reference path tracer (CPU restatement, not Intel "
                      f"ISPC+Embree), -O3 -march=native, {nt} threads",
            "host": cpus, "reference_probe": probe}


def traffic_per_launch(kernel: str, workload: str):
    """Measured HBM bytes per launch of `kernel` on `workload` from the
    committed rocprofv3 counter summary of that workload (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    path = TRAFFIC_FILES.get(workload)
    if path is None or not os.path.exists(path):
        return None
    pat = re.compile(r"^\|\s*" + re.escape(kernel) + r"\s*\|.*\|\s*hbm_bytes_per_launch=(\d+)\s*\|\s*$")
    with open(path) as fh:
        for line in fh:
            m = pat.match(line.strip())
            if m:
                return int(m.group(1))
    return None


def valu_profile(kernel: str):
    """VALU issue and lane activity of `kernel` from the committed rocprofv3
    SQ-counter summary, or None."""
    if not os.path.exists(VALU_FILE):
        return None
    pat = re.compile(r"^\|\s*" + re.escape(kernel) + r"\s*\|.*\|\s*valu=(\S+)\s*\|\s*$")
    with open(VALU_FILE) as fh:
        for line in fh:
            m = pat.match(line.strip())
            if m:
                return dict(kv.split(":", 1) for kv in m.group(1).split(","))
    return None


def bench_c2(args) -> int:
    """64 views of the 1080p scene (scenes.c2_views: seeded positions on a
    radius-30 sphere, resolutions {256^2, 512^2, 768x432, 1024^2}, spp
    {1, 2, 4, 8}, maxPathLength 8) rendered as one batch: the autotuner
    picks the resident grids and the path budget, then the batch is
    captured once into a HIP graph and relaunched --steps times.  value =
    samples of the batch / mean wall time of a launch (host launch + GPU +
    wait); per-view latency = GPU time from the batch start to the view's
    end, p50/p99 over the views of the last launch."""
    import numpy as np

    from ospray_amd import api as A
    from ospray_amd import capi as K
    from ospray_amd import scenes

    sc = scenes.c1()
    views = scenes.c2_views()
    A.init("hip", **args.dparams)
    name, _ = A.device_info()
    s = scenes.to_ospray(sc)
    fbs, rens, cams = [], [], []
    for v in views:
        cams.append(A.new("camera", "perspective", dict(v["camera"])))
        rens.append(A.new("renderer", "pathtracer", {"pixelSamples": v["spp"], "maxPathLength": 8}))
        fbs.append(A.framebuffer(v["width"], v["height"], K.OSP_FB_RGBA32F, K.OSP_FB_COLOR))
    samples = sum(v["width"] * v["height"] * v["spp"] for v in views)
    log(f"{len(views)} views, {samples} samples per batch")
    t = time.perf_counter()
    best_ms, best = A.tune_batch(fbs, rens, cams, s.world, reps=1)
    log(f"autotuner: {time.perf_counter() - t:.1f} s, best batch {best_ms:.2f} ms with traceBlocksPerCU {best[0]}, "
        f"shadeBlocksPerCU {best[1]}, pathBudgetMB {best[2]}, batchStreams {best[3]}, tileWidth {best[4]}")
    t = time.perf_counter()
    A.render_batch(fbs, rens, cams, s.world, graph=True)  # one eager frame, then the capture
    log(f"eager frame and graph capture in {time.perf_counter() - t:.2f} s")
    for _ in range(args.warmup):
        A.render_batch(fbs, rens, cams, s.world, graph=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ms = A.render_batch(fbs, rens, cams, s.world, graph=True)
    elapsed = time.perf_counter() - t0
    eager_t = time.perf_counter()
    A.render_batch(fbs, rens, cams, s.world, graph=False)
    eager_s = time.perf_counter() - eager_t
    lat = np.array(ms)
    img = A.map_color(fbs[0], views[0]["width"], views[0]["height"], K.OSP_FB_RGBA32F)
    if not np.isfinite(img).all():
        raise RuntimeError("non-finite pixels")
    out = {"metric": "batched multi-view path samples/s (64 views of the 1080p scene, one graph launch)",
           "value": round(samples * args.steps / elapsed, 1), "unit": "samples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "C2: 64 views of the C1 scene, seeded positions on a radius-30 sphere, "
                                  "resolutions {256^2, 512^2, 768x432, 1024^2}, spp {1,2,4,8}, maxPathLength 8",
                      "samples_per_batch": samples, "device": name,
                      **({"device_params": args.dparams} if args.dparams else {})},
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
           "autotuner": {"traceBlocksPerCU": best[0], "shadeBlocksPerCU": best[1], "pathBudgetMB": best[2],
                         "batchStreams": best[3], "tileWidth": best[4],
                         "best_eager_batch_ms": round(best_ms, 3)},
           "eager_batch_ms": round(eager_s * 1e3, 3)}
    print(json.dumps(out), flush=True)
    return 0


def measure_c3(args, width: int, height: int) -> dict:
    """Plain vs radiance-cache rendering of the 1080p scene, 1 spp per frame,
    each accumulated over --c3-frames frames and compared (RMSE over RGB)
    with a plain render of --c3-ref-spp samples per pixel."""
    from ospray_amd import scenes

    sc = scenes.c1(width=width, height=height, spp=1)
    return measure_cache(args, sc, args.c3_frames, "C1 scene")


def measure_cache(args, sc, frames: int, scene_name: str) -> dict:
    """Plain vs radiance-cache rendering of the Scene `sc` at its spp per
    frame, each accumulated over `frames` frames and compared (RMSE over
    RGB) with a plain render of --c3-ref-spp samples per pixel (cache after
    bounce --c3-bounce, --c3-train-paths training paths per frame in
    --c3-train-steps Adam steps).  Frame times are ospRenderFrame + ospWait
    wall times (the cache's inference and training included); framebuffer
    maps for the RMSE checkpoints are outside them.  value = (plain time to
    its final RMSE) / (cache time to the same RMSE).  Expects the hip device
    to be current (with profileKernels set); from here on its timers cover
    the cache's inference and training launches only, so the frame times
    carry no event records of the path-tracing stages."""
    import math

    import numpy as np

    from ospray_amd import api as A
    from ospray_amd import capi as K
    from ospray_amd import scenes

    w, h = sc.width, sc.height
    spp_frame = int(sc.renderer.get("pixelSamples", 1))
    name, _ = A.device_info()
    L = K.lib()
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "profileStages", "nrc_infer,nrc_train")
    L.ospDeviceCommit(dev)

    def frame(s, fb):
        f = L.ospRenderFrame(fb.h, s.renderer.h, s.camera.h, s.world.h)
        if not f:
            A.check()
            raise RuntimeError("ospRenderFrame failed")
        L.ospWait(f, K.OSP_TASK_FINISHED)
        L.ospRelease(f)

    t = time.perf_counter()
    ref_spp_frame = 16
    sref = scenes.to_ospray(sc, {"pixelSamples": ref_spp_frame})
    fbr = A.framebuffer(w, h, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_ACCUM)
    nref = max(1, args.c3_ref_spp // ref_spp_frame)
    for i in range(nref):
        frame(sref, fbr)
        if i % 64 == 63:
            log(f"cache comparison: reference {ref_spp_frame * (i + 1)} spp")
    ref = A.map_color(fbr, w, h, K.OSP_FB_RGBA32F)[..., :3].astype(np.float64)
    ref_s = time.perf_counter() - t
    log(f"cache comparison: reference {nref * ref_spp_frame} spp in {ref_s:.1f} s")
    del fbr, sref

    checkpoints = sorted({1, 2, 4, 8, 16, 32, 64, 128, 192, 256, frames} & set(range(1, frames + 1)))

    def run(rp, label):
        s = scenes.to_ospray(sc, rp)
        fb = A.framebuffer(w, h, K.OSP_FB_RGBA32F, K.OSP_FB_COLOR | K.OSP_FB_ACCUM)
        frame(s, fb)  # warm-up frame (first launch of each kernel), then reset
        L.ospResetAccumulation(fb.h)
        A.reset_kernel_stats()
        curve, elapsed = [], 0.0
        for i in range(1, frames + 1):
            t0 = time.perf_counter()
            frame(s, fb)
            elapsed += time.perf_counter() - t0
            if i in checkpoints:
                img = A.map_color(fb, w, h, K.OSP_FB_RGBA32F)[..., :3].astype(np.float64)
                curve.append((i, elapsed, float(np.sqrt(np.mean((img - ref) ** 2)))))
        log(f"cache comparison, {label}: {frames} frames {elapsed:.3f} s rmse {curve[-1][2]:.5f}")
        ks = {k: round(v[1] / frames, 3) for k, v in sorted(A.kernel_stats().items())}
        return curve, s, ks

    plain, _, kplain = run({}, "plain")
    nrc, snrc, knrc = run({"nrc": True, "nrcBounce": args.c3_bounce, "nrcTrainPaths": args.c3_train_paths,
                           "nrcTrainSteps": args.c3_train_steps}, "cache")
    loss, steps, queries, records = A.nrc_state(snrc.renderer)
    target = plain[-1][2]

    def time_to(curve, rmse):
        prev = (0, 0.0, float("inf"))
        for c in curve:
            if c[2] <= rmse:
                if prev[2] == float("inf"):
                    return c[1]
                # log-log interpolation between checkpoints (RMSE ~ t^-1/2)
                f = (math.log(prev[2]) - math.log(rmse)) / max(math.log(prev[2]) - math.log(c[2]), 1e-12)
                return prev[1] * (c[1] / prev[1]) ** f
            prev = c
        return None

    def rmse_at(curve, t):
        # log-log interpolation between checkpoints; None outside the curve
        for (f0, t0, e0), (f1, t1, e1) in zip(curve, curve[1:]):
            if t0 <= t <= t1:
                a = (math.log(t) - math.log(t0)) / max(math.log(t1) - math.log(t0), 1e-12)
                return math.exp(math.log(e0) + a * (math.log(e1) - math.log(e0)))
        return None

    tp, tn = time_to(plain, target), time_to(nrc, target)
    # the two methods' RMSE after the same GPU time (the shorter run's)
    t_eq = min(plain[-1][1], nrc[-1][1])
    equal_time = {"seconds": round(t_eq, 4), "plain_rmse": rmse_at(plain, t_eq), "cache_rmse": rmse_at(nrc, t_eq)}
    return {"time_to_equal_rmse": round(tp / tn, 3) if tn else None, "unit": "x (plain time / cache time)",
            "equal_time": equal_time,
            "workload": f"{scene_name} {w}x{h}, {spp_frame} spp per frame x {frames} frames, cache after bounce "
                        f"{args.c3_bounce}, {args.c3_train_paths} training paths per frame in "
                        f"{args.c3_train_steps} Adam steps (lr 1e-2); reference {nref * ref_spp_frame} spp plain",
            "device": name, "target_rmse": target,
            "cache_bias_floor_rmse": nrc[-1][2],
            "plain": {"rmse": [c[2] for c in plain], "seconds": [round(c[1], 4) for c in plain],
                      "frames": [c[0] for c in plain], "ms_per_frame": round(plain[-1][1] / frames * 1e3, 3),
                      "kernels_ms_per_frame": kplain},
            "cache": {"rmse": [c[2] for c in nrc], "seconds": [round(c[1], 4) for c in nrc],
                      "frames": [c[0] for c in nrc], "ms_per_frame": round(nrc[-1][1] / frames * 1e3, 3),
                      "kernels_ms_per_frame": knrc, "final_training_loss": loss, "adam_steps": steps,
                      "queries_last_frame": queries, "training_records_last_frame": records},
            "time_to_target_s": {"plain": tp, "cache": tn}, "reference_seconds": round(ref_s, 2)}


def bench_c3(args) -> int:
    """--workload c3: the plain vs radiance-cache comparison alone."""
    from ospray_amd import api as A

    A.init("hip", **{"profileKernels": True, **args.dparams})
    r = measure_c3(args, args.width or 1920, args.height or 1080)
    out = {"metric": "plain vs neural-cache time to equal RMSE (1080p scene, 1 spp frames)",
           "value": r["time_to_equal_rmse"], "unit": r["unit"], "n_gpus": 1, "steps": args.c3_frames, "warmup": 1,
           "ms_per_step": r["cache"]["ms_per_frame"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 path tracer, bf16 MLP (fp32 accumulate, fp32 Adam)",
           "data": "synthetic", "config": {"workload": r["workload"], "device": r["device"]}, **r}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample size (0: skip)")
    ap.add_argument("--workload", choices=("c1", "c2", "c3", "c4"), default="c1",
                    help="c1: the 1080p surface scene (default); c2: 64 views of it in one batched graph "
                         "launch; c3: plain vs radiance-cache time to equal RMSE on it; c4: surfaces inside a "
                         "512^3 volume")
    ap.add_argument("--c3-frames", type=int, default=256, help="c3: 1-spp frames per method")
    ap.add_argument("--c3-ref-spp", type=int, default=4096, help="c3: samples per pixel of the plain reference")
    ap.add_argument("--c3-bounce", type=int, default=2, help="c3: depth of the cache vertex (renderer nrcBounce)")
    ap.add_argument("--c3-train-paths", type=int, default=65536,
                    help="c3: training paths per frame (nrcTrainPaths)")
    ap.add_argument("--c3-train-steps", type=int, default=4, help="c3: Adam steps per frame (nrcTrainSteps)")
    ap.add_argument("--no-c3", action="store_true",
                    help="c1, c4: skip the plain vs radiance-cache leg (time to equal RMSE)")
    ap.add_argument("--c4-cache-frames", type=int, default=256,
                    help="c4: frames (of the scene's spp) per method in the radiance-cache leg")
    ap.add_argument("--width", type=int, default=None)
    ap.add_argument("--height", type=int, default=None)
    ap.add_argument("--spp", type=int, default=None)
    ap.add_argument("--device-param", action="append", default=[], metavar="KEY=VALUE",
                    help="integer or boolean (true/false) parameter of the hip device (traceBlocksPerCU, "
                         "shadeBlocksPerCU, pathBudgetMB, traversalStats); repeatable")
    ap.add_argument("--all-stage-timers", action="store_true",
                    help="time every stage of the frame with HIP events (default: only the kernels the roofline "
                         "line considers; each timed launch adds two event records to the stream)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only time the plain copy kernel on BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the frame would be split by "
              "tiles or sample ranges across devices; that path is not built)", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    from ospray_amd import api as A
    from ospray_amd import capi as K
    from ospray_amd import scenes

    args.dparams = {}
    for kv in args.device_param:
        k, _, v = kv.partition("=")
        if v.lower() in ("true", "false"):
            args.dparams[k] = v.lower() == "true"
        else:
            try:
                args.dparams[k] = int(v)
            except ValueError:
                args.dparams[k] = v
    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c2":
        return bench_c2(args)

    if args.copy_calibration:
        A.init("hip")
        ms = A.copy_bandwidth(args.copy_calibration, 10)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms": ms,
                          "gbs": 2 * args.copy_calibration / (ms * 1e-3) / 1e9}), flush=True)
        return 0

    t = time.perf_counter()
    if args.workload == "c4":
        args.width, args.height, args.spp = args.width or 1280, args.height or 720, args.spp or 2
        sc = scenes.c4(width=args.width, height=args.height, spp=args.spp)
    else:
        args.width, args.height, args.spp = args.width or 1920, args.height or 1080, args.spp or 4
        sc = scenes.c1(width=args.width, height=args.height, spp=args.spp)
    ntris = sc.num_triangles
    log(f"scene generated in {time.perf_counter() - t:.1f} s ({ntris} instanced triangles)")

    dparams = args.dparams
    # the timed frames carry HIP event timers around the roofline candidates
    # only (the traversal and delta-tracking kernels): timers around every
    # stage add about 0.9 ms of event records per frame
    # (the transparent-shadow pass, a few rays per bounce, is never dominant)
    stages = "extend,shadow" + (",volume,resolve" if args.workload == "c4" else "")
    A.init("hip", **{"profileKernels": True, "profileStages": "" if args.all_stage_timers else stages, **dparams})
    name, cus = A.device_info()
    log(f"device {name}, {cus} CUs")
    t = time.perf_counter()
    s = scenes.to_ospray(sc)
    log(f"scene committed (host BVH build + upload) in {time.perf_counter() - t:.1f} s")
    ch = K.OSP_FB_COLOR | K.OSP_FB_ACCUM | K.OSP_FB_VARIANCE | K.OSP_FB_DEPTH
    fb = A.framebuffer(sc.width, sc.height, K.OSP_FB_RGBA32F, ch)
    L = K.lib()

    def frame():
        f = L.ospRenderFrame(fb.h, s.renderer.h, s.camera.h, s.world.h)
        if not f:
            A.check()
            raise RuntimeError("ospRenderFrame failed")
        L.ospWait(f, K.OSP_TASK_FINISHED)
        L.ospRelease(f)

    for i in range(args.warmup):
        frame()
        log(f"warmup {i + 1}/{args.warmup}")
    A.check()
    A.reset_kernel_stats()
    t0 = time.perf_counter()
    for i in range(args.steps):
        frame()
    elapsed = time.perf_counter() - t0
    A.check()
    kstats_raw = A.kernel_stats()
    # with the profileBounces device parameter the stage timers are kept per
    # bounce ("extend@3"): each stage's sum over the bounces is its time
    kstats = {}
    for k, (launches, ms) in kstats_raw.items():
        stage = k.split("@")[0]
        l0, m0 = kstats.get(stage, (0, 0.0))
        kstats[stage] = (l0 + launches, m0 + ms)
    log(f"timed {args.steps} frames in {elapsed:.3f} s")

    samples = sc.width * sc.height * args.spp
    value = samples * args.steps / elapsed

    # traversal counts of one more frame (not timed)
    dev = L.ospGetCurrentDevice()
    A.set_device_param(dev, "traversalStats", True)
    L.ospDeviceCommit(dev)
    frame()
    tr = A.traversal_stats()
    vst = A.volume_stats()
    A.set_device_param(dev, "traversalStats", False)
    L.ospDeviceCommit(dev)
    col = A.map_color(fb, sc.width, sc.height, K.OSP_FB_RGBA32F)
    import numpy as np

    if not np.isfinite(col).all():
        raise RuntimeError("non-finite pixels in the rendered frame")
    log(f"traversal counters: {tr}")

    kern = {}
    for k, key in (("extend", 0), ("shadow", 3), ("shadow_trans", 6)):
        nodes, prims, rays = tr[key:key + 3]
        launches, ms = kstats.get(k, (0, 0.0))
        if not launches:
            continue
        per_frame_bytes = NODE_BYTES * nodes + PRIM_BYTES * prims + RAY_BYTES * rays
        record_bytes = NODE_RECORD_BYTES * nodes + PRIM_BYTES * prims + RAY_BYTES * rays
        ms_frame = ms / args.steps
        kern[k] = {"ms_per_frame": ms_frame, "launches_per_frame": launches / args.steps,
                   "bytes_per_frame": per_frame_bytes, "rays_per_frame": rays,
                   "nodes_per_ray": nodes / max(rays, 1), "prims_per_ray": prims / max(rays, 1),
                   "gbs": per_frame_bytes / (ms_frame * 1e-3) / 1e9,
                   "record_gbs": record_bytes / (ms_frame * 1e-3) / 1e9}
    if kstats.get("volume", (0, 0.0))[0]:
        launches, ms = kstats["volume"]
        samples_v, cells_v, tracks_v = vst[:3]
        per_frame_bytes = SAMPLE_BYTES * samples_v + CELL_BYTES * cells_v + VOL_RAY_BYTES * tr[2]
        ms_frame = ms / args.steps
        kern["volume"] = {"ms_per_frame": ms_frame, "launches_per_frame": launches / args.steps,
                          "bytes_per_frame": per_frame_bytes, "rays_per_frame": 0,
                          "samples_per_frame": samples_v, "cells_per_frame": cells_v,
                          "tracks_per_frame": tracks_v, "gbs": per_frame_bytes / (ms_frame * 1e-3) / 1e9}
        log(f"volume counters: free path {vst[:3]}, transmittance {vst[3:]}")
    if kstats.get("resolve", (0, 0.0))[0] and vst[5]:
        # with volumes k_resolve runs the shadow records' transmittance estimates
        launches, ms = kstats["resolve"]
        samples_t, cells_t, tracks_t = vst[3:6]
        per_frame_bytes = SAMPLE_BYTES * samples_t + CELL_BYTES * cells_t + TRACK_BYTES * tracks_t
        ms_frame = ms / args.steps
        kern["resolve"] = {"ms_per_frame": ms_frame, "launches_per_frame": launches / args.steps,
                           "bytes_per_frame": per_frame_bytes, "rays_per_frame": 0,
                           "samples_per_frame": samples_t, "cells_per_frame": cells_t,
                           "tracks_per_frame": tracks_t, "gbs": per_frame_bytes / (ms_frame * 1e-3) / 1e9}
    dom = max(kern, key=lambda k: kern[k]["ms_per_frame"])
    d = kern[dom]
    lpf = d["launches_per_frame"]
    traffic = traffic_per_launch(f"k_{dom}", args.workload)
    launch_ms = d["ms_per_frame"] / lpf
    frac_alg = d["gbs"] / HBM_PEAK_GBS
    # measured DRAM bytes per launch (committed counter summary) over the
    # live launch time; the HBM fraction stated is the lower of the two
    frac_dram = traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None
    roofline = {"bound": "hbm", "achieved": round(d["gbs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(min(frac_alg, frac_dram) if frac_dram is not None else frac_alg, 4),
                "traffic": traffic, "kernel": f"k_{dom}",
                "frac_alg": round(frac_alg, 4),
                "frac_dram": round(frac_dram, 4) if frac_dram is not None else None,
                "alg_bytes": (f"{SAMPLE_BYTES} B per voxel sample + {CELL_BYTES} B per macrocell interval + "
                              f"{VOL_RAY_BYTES} B per queued ray") if dom == "volume" else
                             (f"{SAMPLE_BYTES} B per voxel sample + {CELL_BYTES} B per macrocell interval + "
                              f"{TRACK_BYTES} B per transmittance estimate") if dom == "resolve" else
                             (f"{NODE_BYTES} B per node visit + {PRIM_BYTES} B per primitive test + {RAY_BYTES} B "
                              f"per ray"),
                "bytes_per_launch": int(d["bytes_per_frame"] / lpf),
                "launch_ms": round(launch_ms, 4)}
    if "record_gbs" in d:
        roofline["frac_alg_128B_node_record"] = round(d["record_gbs"] / HBM_PEAK_GBS, 4)
    valu = valu_profile(f"k_{dom}") if args.workload == "c1" else None
    if valu:
        roofline["valu"] = valu
    rays_total = sum(kern[k]["rays_per_frame"] for k in kern)
    copy_ms = A.copy_bandwidth(1 << 30, 10)
    roofline["copy_kernel_gbs"] = round(2 * (1 << 30) / (copy_ms * 1e-3) / 1e9, 1)

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(sc, args.cpu_seconds)

    ms_step = elapsed / args.steps * 1e3
    mpl = sc.renderer["maxPathLength"]
    if args.workload == "c4":
        workload = (f"surface+volume pathtracer {sc.width}x{sc.height} {args.spp}spp maxPathLength {mpl}, "
                    "C0 spheres + ground in a 512^3 fBm density volume (8-point transfer function, opacity "
                    "scale 0.5), delta tracking, quad + ambient light")
    else:
        workload = (f"plain pathtracer {sc.width}x{sc.height} {args.spp}spp maxPathLength {mpl}, "
                    "fBm heightfield 1448^2 (4,187,618 tris) + 256 icosphere instances, sun + 4 quad lights")
    out = {"metric": METRIC if args.workload == "c1" else "path samples/s, surface+volume scene",
           "value": round(value, 1), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": workload, "width": sc.width, "height": sc.height, "spp": args.spp,
                      "max_path_length": mpl, "instanced_triangles": int(ntris), "device": name,
                      **({"device_params": dparams} if dparams else {})},
           "rays_per_s": round(rays_total * args.steps / elapsed, 1),
           "kernels_ms_per_frame": {k: round(v[1] / args.steps, 3) for k, v in sorted(kstats.items())},
           **({"stages_ms_per_frame": {k: round(v[1] / args.steps, 3) for k, v in sorted(kstats_raw.items())}}
              if kstats_raw.keys() != kstats.keys() else {}),
           "traversal": {k: {"nodes_per_ray": round(tr[key] / max(tr[key + 2], 1), 2),
                             "prims_per_ray": round(tr[key + 1] / max(tr[key + 2], 1), 2),
                             "rays_per_frame": tr[key + 2],
                             **({"gbs": round(kern[k]["gbs"], 1)} if k in kern else {})}
                         for k, key in (("extend", 0), ("shadow", 3), ("shadow_trans", 6))},
           "roofline": roofline, "cpu_baseline": cpu}
    if args.workload == "c1" and not args.no_c3:
        out["c3"] = measure_c3(args, 1920, 1080)
        log(f"c3: time to equal RMSE {out['c3']['time_to_equal_rmse']}x")
    if args.workload == "c4" and not args.no_c3:
        # the cache at surface and volume scattering vertices, against the
        # plain render of the same scene
        out["c4_cache"] = measure_cache(args, sc, args.c4_cache_frames, "surface+volume scene")
        et = out["c4_cache"]["equal_time"]
        log(f"c4 cache: time to equal RMSE {out['c4_cache']['time_to_equal_rmse']}x; after {et['seconds']} s "
            f"RMSE plain {et['plain_rmse']}, cache {et['cache_rmse']}")
    if "volume" in kern:
        v = kern["volume"]
        out["volume"] = {"free_path": {"samples": vst[0], "cells": vst[1], "tracks": vst[2],
                                       "ms_per_frame": round(v["ms_per_frame"], 3), "gbs": round(v["gbs"], 1)},
                         "transmittance": {"samples": vst[3], "cells": vst[4], "tracks": vst[5]}}
        if "resolve" in kern:
            out["volume"]["transmittance"].update(ms_per_frame=round(kern["resolve"]["ms_per_frame"], 3),
                                                  gbs=round(kern["resolve"]["gbs"], 1))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
