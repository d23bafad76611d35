"""Builds the native libraries in-tree.

  ospray_amd/lib/libospray.so             C API front end (g++)
  ospray_amd/lib/libospray_module_hip.so  the "hip" device: gfx950 kernels
                                          and host code (hipcc)

Run `python -m ospray_amd.build` (or __graft_entry__.build()).  Object files
go to ospray_amd/_build/.  Every HIP translation unit is compiled with
-ffp-contract=off and IEEE division/sqrt so device float results are
reproducible on the CPU, and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ospray_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("OSPRAY_HIP_ARCH", "gfx950")

# the core exports the ospcore C++ classes that device modules derive from,
# so its symbols keep default visibility
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            f"-I{INCLUDE}"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-fvisibility=hidden", f"-I{INCLUDE}", f"-I{os.path.join(CSRC, 'hip')}",
            "-Wno-unused-command-line-argument"]

CORE_SRC = ["core/managed.cpp", "core/api.cpp", "core/util.cpp"]
HIP_CXX_SRC = ["hip/bvh_build.cpp"]
HIP_SRC = ["hip/pt_kernels.hip", "hip/hip_device.hip", "hip/nrc.hip", "hip/bvh_device.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIPFLAGS, "-c", path, "-o", obj]
    else:
        cmd = ["g++", *CXXFLAGS, f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)
    srcs = CORE_SRC + HIP_CXX_SRC + HIP_SRC
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_compile, srcs))
    core = os.path.join(LIB, "libospray.so")
    _run(["g++", "-shared", "-o", core + ".tmp", *[_obj(s) for s in CORE_SRC], "-ldl", "-lpthread",
          "-Wl,-soname,libospray.so"])
    os.replace(core + ".tmp", core)
    mod = os.path.join(LIB, "libospray_module_hip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", mod + ".tmp",
          *[_obj(s) for s in HIP_CXX_SRC + HIP_SRC], f"-L{LIB}", "-lospray", "-Wl,-rpath,$ORIGIN",
          "-Wl,-soname,libospray_module_hip.so", "-lpthread"])
    os.replace(mod + ".tmp", mod)
    if verbose:
        print(f"built {core}\nbuilt {mod}")


def build_oracle(verbose: bool = True) -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])
    if verbose:
        print(f"built {os.path.join(ROOT, 'oracle', 'liborc.so')}")


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
