"""Build recipe of the package's native libraries (run by __graft_entry__.build()).

libvkgv_hip.so   hipcc for gfx950, all HIP translation units in one shared library
libvkgv_host.so  g++ (host producers: camera, World, meshlets, scenes)
vkgv_diff.so     LibTorch extension of the differentiable form (csrc/diff/), built
                 by torch.utils.cpp_extension with hipcc for gfx950

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

HIP_SOURCES = ["api.hip", "cull.hip", "raster.hip", "hiz.hip", "resolve.hip", "grad.hip", "anim.hip", "calib.hip"]
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


# AddressSanitizer + UndefinedBehaviorSanitizer build of the host library (CPU
# code only), for the sanitized test run (tests/test_sanitizers.py); loaded in
# place of the plain build through VKGV_HOST_LIB.
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=undefined"]


def build_host(verbose=True, sanitize=False):
    out_dir = os.path.join(LIB, "sanitized") if sanitize else LIB
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libvkgv_host.so")
    opt = SANITIZE_FLAGS if sanitize else ["-O2"]
    _run(["g++", *opt, "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-Wextra", "-Wno-unused-parameter", "-I", INCLUDE, "-o", out,
          os.path.join(CSRC, "host", "host.cpp"), os.path.join(CSRC, "host", "meshopt.cpp")])
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


def build_diff(verbose=False):
    """The differentiable form: its kernels (csrc/diff/diff.hip) with hipcc into
    lib/libvkgv_diffk.so, and the LibTorch binding (csrc/diff/diff_ext.cpp, host
    C++ only) with torch.utils.cpp_extension into lib/vkgv_diff.so, linked to
    the kernel library (resolved by preloading it with RTLD_GLOBAL before the
    binding is imported)."""
    import shutil
    os.makedirs(LIB, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    kout = os.path.join(LIB, "libvkgv_diffk.so")
    _run([hipcc, *HIPCC_FLAGS, "-Wl,-soname,libvkgv_diffk.so", "-I", os.path.join(CSRC, "diff"), "-o", kout + ".tmp",
          os.path.join(CSRC, "diff", "diff.hip")])
    os.replace(kout + ".tmp", kout)
    import ctypes
    import torch
    # the binding's NEEDED libvkgv_diffk.so resolves to this already-loaded library
    ctypes.CDLL(kout, mode=ctypes.RTLD_GLOBAL)
    from torch.utils import cpp_extension
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    bdir = os.path.join(LIB, "_build_diff")
    os.makedirs(bdir, exist_ok=True)
    mod = cpp_extension.load(
        name="vkgv_diff", sources=[os.path.join(CSRC, "diff", "diff_ext.cpp")],
        extra_cflags=["-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-I/opt/rocm/include",
                      "-I" + os.path.join(CSRC, "diff")],
        extra_ldflags=["-L" + LIB, "-lvkgv_diffk", "-L/opt/rocm/lib", "-lamdhip64",
                       "-L" + torch_lib, "-lc10_hip"],
        with_cuda=False, build_directory=bdir, verbose=verbose, is_python_module=True)
    out = os.path.join(LIB, "vkgv_diff.so")
    shutil.copyfile(mod.__file__, out + ".tmp")
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_host()
    build_hip()
    build_diff()


if __name__ == "__main__":
    what = sys.argv[1:] or ["host", "hip", "diff"]
    if "host" in what:
        build_host()
    if "hip" in what:
        build_hip()
    if "diff" in what:
        build_diff()
