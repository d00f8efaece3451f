"""The CPU code under AddressSanitizer and UndefinedBehaviorSanitizer.

The oracle (make sanitize) and the host library (meshlet builder, camera,
World traversal, scene generators and the EXT_meshopt_compression decoders,
which parse untrusted bitstreams) are rebuilt with
-fsanitize=address,undefined, and the CPU tests that exercise them run in a
child process that loads those builds (VKGV_ORACLE_LIB, VKGV_HOST_LIB) with the
sanitizer runtimes preloaded.  Any out-of-bounds access, use after free or
undefined behaviour aborts the child and fails this test.  The mutated-stream
test of the decoders passes each stream in a malloc'd buffer of its exact size
(see meshopt_exact_buffers below), so that a read past its end is caught.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FILES = ["tests/test_meshopt.py", "tests/test_host.py", "tests/test_gltf.py", "tests/test_oracle_raster.py",
         "tests/test_oracle_geometry.py", "tests/test_oracle_texture.py", "tests/test_oracle_anim.py"]


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_cpu_code_under_address_and_undefined_behaviour_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("the gcc sanitizer runtimes are not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    sys.path.insert(0, ROOT)
    from vk_gltf_viewer_amd import build as B
    host = B.build_host(sanitize=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan}:{ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:allocator_may_return_null=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
        "VKGV_HOST_LIB": host,
        "VKGV_ORACLE_LIB": os.path.join(ROOT, "oracle", "build", "sanitized", "libvkgv_oracle.so"),
        "VKGV_MESHOPT_EXACT_BUFFERS": "1",
        "OMP_NUM_THREADS": "2",
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-n", "4", *FILES], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout[-4000:] + "\n" + r.stderr[-4000:])
    print(tail)
    assert r.returncode == 0, tail
    assert "passed" in r.stdout
