"""The oracle's pinning tests reject plausible mistakes.

Each planted error k (ORC_PLANT in oracle/refcpu/orc_render.cpp) is a
one-line mistake in one function that the closed-form and distribution
tests of test_oracle_pins.py pin; the oracle is built with it
(`make -C oracle planted`) and the tests named for it must fail, while the
unplanted oracle passes them all (the same tests run in the suite):

  1  camera: the full vertical field of view where half belongs
     (PerspectiveCamera.cpp:30-110);
  2  Gaussian pixel filter: sigma = width / 4 instead of width / 6
     (pf/PixelFilter.cpp:55-56);
  3  Variance_kernel: the per-pixel difference not normalised by
     sqrt(r + g + b + 1 - a) (fb/frame_ops/Variance.ispc:11-59);
  4  light selection: the selection weight without the sample count
     (NextEventEstimation.ispc:54-63);
  5  sun cone: cos(theta) drawn from a non-uniform density
     (math/sampling.ih:92-97);
  6  glass transparency: the Fresnel term at normal incidence whatever the
     angle (materials/Glass.ispc:37-52).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")

PLANTS = {
    1: ["test_camera_rays_match_fp64_pinhole", "test_camera_field_of_view"],
    2: ["test_gaussian_pixel_filter_distribution"],
    3: ["test_variance_tasks_hand_computed", "test_variance_random_frame_fp64"],
    4: ["test_five_lights_on_plane"],
    5: ["test_sun_cone_directions", "test_cone_sun_on_plane"],
    6: ["test_sun_through_glass_slab"],
}


@pytest.fixture(scope="module")
def planted_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("planted")
    procs, libs = [], {}
    for k in PLANTS:
        libs[k] = str(out / f"liborc_plant{k}.so")
        procs.append(subprocess.Popen(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "planted", f"PLANT={k}",
                                       f"OUT={libs[k]}"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out_text, _ = p.communicate(timeout=600)
        assert p.returncode == 0, out_text.decode()
    return libs


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pins_reject_planted_error(planted_libs, plant):
    names = PLANTS[plant]
    env = dict(os.environ, ORC_LIB=planted_libs[plant])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", PINS, "-k", " or ".join(names)],
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-3000:]
    failed = {line.split("::")[1].split("[")[0].split(" ")[0] for line in r.stdout.splitlines()
              if line.startswith("FAILED ")}
    assert set(names) <= failed, (sorted(failed), r.stdout[-3000:])
