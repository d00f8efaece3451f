"""EXT_meshopt_compression buffer-view decoding through libvkgv_host.so
(csrc/host/meshopt.cpp), as the reference's CompressedBufferDataAdapter
does it (assets.cpp:111-171), but with the decoder's status checked."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N

MODES = {"ATTRIBUTES": 0, "TRIANGLES": 1, "INDICES": 2}
FILTERS = {"NONE": 0, "OCTAHEDRAL": 1, "QUATERNION": 2, "EXPONENTIAL": 3}
_ERRORS = {-1: "bad header, version or arguments", -2: "stream ends early", -3: "trailing data"}


class MeshoptError(ValueError):
    pass


def decode(data: bytes, count: int, stride: int, mode: str | int = "ATTRIBUTES", filter: str | int = "NONE") -> bytes:
    """Decodes one compressed buffer view to count * stride bytes."""
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    f = FILTERS[filter] if isinstance(filter, str) else int(filter)
    src = np.frombuffer(bytes(data), np.uint8)
    out = np.zeros(max(1, count * stride), np.uint8)
    rc = N.host().vkgv_meshopt_decode(out.ctypes.data, count, stride, src.ctypes.data if src.size else None,
                                      src.size, m, f)
    if rc != 0:
        raise MeshoptError(f"meshopt decode failed ({rc}: {_ERRORS.get(rc, 'error')}); mode {mode}, filter {filter}, "
                           f"count {count}, stride {stride}")
    return out[:count * stride].tobytes()


def _call(name, data, count, size):
    data = bytes(data)
    if os.environ.get("VKGV_MESHOPT_EXACT_BUFFERS"):
        # the sanitized test run: source and destination in malloc'd blocks of
        # their exact sizes, so that an access past either end is detected
        libc = C.CDLL(None)
        libc.malloc.restype = C.c_void_p
        libc.malloc.argtypes = [C.c_size_t]
        libc.free.argtypes = [C.c_void_p]
        n_out = count * size
        src = libc.malloc(max(1, len(data)))
        dst = libc.malloc(max(1, n_out))
        try:
            C.memmove(src, data, len(data))
            C.memset(dst, 0, max(1, n_out))
            rc = getattr(N.host(), name)(dst, count, size, src if data else None, len(data))
            out = np.frombuffer(C.string_at(dst, n_out), np.uint8).copy()
        finally:
            libc.free(src)
            libc.free(dst)
        return rc, out
    src = np.frombuffer(data, np.uint8)
    out = np.zeros(max(1, count * size), np.uint8)
    rc = getattr(N.host(), name)(out.ctypes.data, count, size, src.ctypes.data if src.size else None, src.size)
    return rc, out[:count * size]


def decode_vertex_buffer(data, count, vertex_size):
    """(rc, bytes) of meshopt_decodeVertexBuffer's equivalent."""
    return _call("vkgv_meshopt_decode_vertex_buffer", data, count, vertex_size)


def decode_index_buffer(data, count, index_size):
    return _call("vkgv_meshopt_decode_index_buffer", data, count, index_size)


def decode_index_sequence(data, count, index_size):
    return _call("vkgv_meshopt_decode_index_sequence", data, count, index_size)


def filter_in_place(name: str, arr: np.ndarray, count: int, stride: int):
    """Runs filter `name` ("oct", "quat", "exp") over a contiguous array."""
    assert arr.flags.c_contiguous and arr.nbytes >= count * stride
    getattr(N.host(), f"vkgv_meshopt_decode_filter_{name}")(arr.ctypes.data, count, stride)
