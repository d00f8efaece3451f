"""Build recipe of the package's native libraries (run by __graft_entry__.build()).

libvkgv_hip.so   hipcc for gfx950, all HIP translation units in one shared library
libvkgv_host.so  g++ (host producers: camera, World, meshlets, scenes)

Both are built in-tree under vk_gltf_viewer_amd/lib/.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["api.hip", "cull.hip", "raster.hip", "hiz.hip", "resolve.hip"]
HIP_HEADERS = ["internal.hpp", "device_math.hpp"]

# -ffp-contract=off: the kernels fix their floating-point evaluation order.
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libvkgv_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INCLUDE, "-o", out,
          os.path.join(CSRC, "host", "host.cpp")])
    return out


def build_hip(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libvkgv_hip.so")
    tmp = out + ".tmp"
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    _run([hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-o", tmp,
          *[os.path.join(CSRC, s) for s in HIP_SOURCES]])
    os.replace(tmp, out)
    return out


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1:] or ["host", "hip"]
    if "host" in what:
        build_host()
    if "hip" in what:
        build_hip()
