"""ctypes bindings of the two native libraries of this package.

``libvkgv_hip.so``  the HIP implementation of the frame (include/vkgv.h)
``libvkgv_host.so`` host-side producers: camera, World, meshlets, scenes
                    (include/vkgv_host.h)

The structures below mirror the C headers field for field; their sizes are
checked against the contract sizes at import time.  Loading fails loudly: there
is no Python or CPU fallback for anything in libvkgv_hip.so.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# VKGV_HIP_LIB selects another build of the HIP library
HIP_LIB = os.environ.get("VKGV_HIP_LIB") or os.path.join(LIB_DIR, "libvkgv_hip.so")
# VKGV_HOST_LIB selects another build of the host library (the sanitized one)
HOST_LIB = os.environ.get("VKGV_HOST_LIB") or os.path.join(LIB_DIR, "libvkgv_host.so")

# ---------------------------------------------------------------------------
# Constants (include/vkgv.h)
MAX_VERTICES = 64
MESHLET_TRIANGLES = 124
TRIANGLE_BITS = 7
MAX_DRAWS = 1 << 25
KEY_CLEAR = 0xFFFFFFFFFFFFFFFF
VISBUFFER_CLEAR = 0xFFFFFFFF

CULL_FRUSTUM = 1 << 0
CULL_OCCLUSION_PREV = 1 << 1
CULL_TWO_PASS = 1 << 2
CULL_CONE = 1 << 3
RESOLVE_BARY = 1 << 4
BUILD_PYRAMID = 1 << 5
TIMING = 1 << 6
NO_BACKFACE = 1 << 7
SKIP_RESOLVE = 1 << 8
GRAPH = 1 << 9
RESOLVE_TEXTURED = 1 << 10
MOTION_VECTORS = 1 << 11
STATS = 1 << 12
INTERP_LINEAR, INTERP_STEP, INTERP_CUBICSPLINE = 0, 1, 2
PATH_TRANSLATION, PATH_ROTATION, PATH_SCALE = 0, 1, 2
WRAP_REPEAT, WRAP_CLAMP_TO_EDGE, WRAP_MIRRORED_REPEAT = 0, 1, 2
MODE_REFERENCE = CULL_FRUSTUM | CULL_OCCLUSION_PREV | BUILD_PYRAMID

OK = 0
ERROR_INVALID_ARGUMENT = -1
ERROR_OUT_OF_MEMORY = -2
ERROR_HIP = -3
ERROR_LIMIT = -4
ERROR_STATE = -5


class Meshlet(C.Structure):
    _fields_ = [("vertex_offset", C.c_uint32), ("triangle_offset", C.c_uint32),
                ("vertex_count", C.c_uint8), ("triangle_count", C.c_uint8), ("pad", C.c_uint8 * 2),
                ("aabb_extents", C.c_float * 3), ("aabb_center", C.c_float * 3)]


class Vertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("color", C.c_uint8 * 4), ("normal", C.c_uint8 * 3),
                ("pad", C.c_uint8), ("uv", C.c_uint16 * 2)]


class MeshletDraw(C.Structure):
    _fields_ = [("primitive_index", C.c_uint32), ("meshlet_index", C.c_uint32), ("transform_index", C.c_uint32)]


class Primitive(C.Structure):
    _fields_ = [("vertex_index_buffer", C.c_uint64), ("primitive_index_buffer", C.c_uint64),
                ("vertex_buffer", C.c_uint64), ("meshlet_buffer", C.c_uint64),
                ("aabb_extents", C.c_float * 3), ("aabb_center", C.c_float * 3),
                ("meshlet_count", C.c_uint32), ("material_index", C.c_uint32)]


class TextureDesc(C.Structure):
    _fields_ = [("rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("wrap_s", C.c_uint32), ("wrap_t", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("albedo_factor", C.c_float * 4), ("albedo_index", C.c_uint32), ("uv_offset", C.c_float * 2),
                ("uv_scale", C.c_float * 2), ("uv_rotation", C.c_float), ("alpha_cutoff", C.c_float),
                ("double_sided", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("prev_view_projection", C.c_float * 16), ("prev_occlusion_view_projection", C.c_float * 16),
                ("view_projection", C.c_float * 16), ("occlusion_view_projection", C.c_float * 16),
                ("frustum", (C.c_float * 4) * 6)]


class MeshletCone(C.Structure):
    _fields_ = [("apex", C.c_float * 3), ("axis", C.c_float * 3), ("cutoff", C.c_float), ("pad", C.c_float)]


class MeshDesc(C.Structure):
    _fields_ = [("vertex_indices", C.c_void_p), ("vertex_index_count", C.c_uint64),
                ("triangle_indices", C.c_void_p), ("triangle_index_bytes", C.c_uint64),
                ("vertices", C.c_void_p), ("vertex_count", C.c_uint64),
                ("meshlets", C.c_void_p), ("meshlet_count", C.c_uint32), ("pad", C.c_uint32),
                ("cones", C.c_void_p)]


class PrimitiveDesc(C.Structure):
    _fields_ = [("mesh_index", C.c_uint32), ("material_index", C.c_uint32),
                ("aabb_extents", C.c_float * 3), ("aabb_center", C.c_float * 3)]


class FrameDesc(C.Structure):
    _fields_ = [("camera", Camera), ("eye", C.c_float * 3), ("flags", C.c_uint32), ("stream", C.c_void_p)]


class FrameStats(C.Structure):
    _fields_ = [("draws", C.c_uint32), ("visible_early", C.c_uint32), ("visible_late", C.c_uint32),
                ("big_triangles", C.c_uint32), ("big_tiles", C.c_uint32), ("clipped_triangles", C.c_uint32),
                ("overflow", C.c_uint32), ("timed_frames", C.c_uint32),
                ("ms_clear", C.c_float), ("ms_cull", C.c_float), ("ms_raster", C.c_float), ("ms_hiz", C.c_float),
                ("ms_cull_late", C.c_float), ("ms_raster_late", C.c_float), ("ms_resolve", C.c_float),
                ("ms_pyramid", C.c_float), ("ms_total", C.c_float), ("max_tile_tris", C.c_uint32),
                ("raster_triangles", C.c_uint64), ("fragments", C.c_uint64), ("fragment_atomics", C.c_uint64),
                ("fragment_stores", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("small_max_extent", C.c_uint32), ("raster_blocks_per_cu", C.c_uint32), ("tile_item_tris", C.c_uint32),
                ("resolve_cache_slots", C.c_uint32)]


class CreateInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("big_triangle_capacity", C.c_uint32), ("clip_triangle_capacity", C.c_uint32),
                ("tile_triangle_capacity", C.c_uint32)]


class ViewDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("orthographic", C.c_uint32), ("pad", C.c_uint32),
                ("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("znear", C.c_float), ("zfar", C.c_float),
                ("ortho_half_width", C.c_float), ("ortho_half_height", C.c_float)]


class Node(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("first_child", C.c_uint32), ("child_count", C.c_uint32), ("pad", C.c_uint32),
                ("translation", C.c_float * 3), ("rotation", C.c_float * 4), ("scale", C.c_float * 3)]


class MeshletBuild(C.Structure):
    _fields_ = [("vertex_indices", C.POINTER(C.c_uint32)), ("vertex_index_count", C.c_uint64),
                ("triangle_indices", C.POINTER(C.c_uint8)), ("triangle_index_bytes", C.c_uint64),
                ("meshlets", C.POINTER(Meshlet)), ("meshlet_count", C.c_uint32),
                ("cones", C.POINTER(MeshletCone))]


# Contract sizes (shaders/mesh_common.h.glsl layouts).
_SIZES = {Meshlet: 36, Vertex: 24, MeshletDraw: 12, Primitive: 64, Material: 48, Camera: 352, MeshletCone: 32}
for _t, _n in _SIZES.items():
    assert C.sizeof(_t) == _n, f"{_t.__name__} is {C.sizeof(_t)} bytes, contract says {_n}"

_p = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_i32 = C.c_int32


def _bind(lib, name, restype, *argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_host = None
_hip = None


def host():
    """libvkgv_host.so (camera, World, meshlets, scenes, meshopt decoders)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: run __graft_entry__.build()")
        lib = C.CDLL(HOST_LIB)
        _bind(lib, "vkgv_camera_from_view", None, C.POINTER(ViewDesc), C.POINTER(Camera))
        _bind(lib, "vkgv_camera_advance", None, C.POINTER(Camera), C.POINTER(ViewDesc), C.POINTER(Camera))
        _bind(lib, "vkgv_build_meshlets", C.c_int, _p, _u32, _p, _u64, _u32, _u32, C.POINTER(MeshletBuild))
        _bind(lib, "vkgv_meshlet_build_free", None, C.POINTER(MeshletBuild))
        _bind(lib, "vkgv_scene_generate", _p, _u32, _u64, _u32)
        _bind(lib, "vkgv_scene_free", None, _p)
        _bind(lib, "vkgv_scene_error", C.c_char_p)
        for nm in ("mesh", "primitive", "material", "draw", "transform", "view"):
            _bind(lib, f"vkgv_scene_{nm}_count", _u32, _p)
        _bind(lib, "vkgv_scene_meshes", C.POINTER(MeshDesc), _p)
        _bind(lib, "vkgv_scene_primitives", C.POINTER(PrimitiveDesc), _p)
        _bind(lib, "vkgv_scene_materials", C.POINTER(Material), _p)
        _bind(lib, "vkgv_scene_draws", C.POINTER(MeshletDraw), _p)
        _bind(lib, "vkgv_scene_transforms", C.POINTER(C.c_float), _p)
        _bind(lib, "vkgv_scene_views", C.POINTER(ViewDesc), _p)
        _bind(lib, "vkgv_scene_triangle_count", _u64, _p)
        _bind(lib, "vkgv_world_build", C.c_int, C.POINTER(Node), _u32, _p, _p, _u32, _p, _p, _p,
              _p, C.POINTER(_u32), _p, C.POINTER(_u32))
        for nm in ("vertex_buffer", "index_buffer", "index_sequence"):
            _bind(lib, f"vkgv_meshopt_decode_{nm}", C.c_int, _p, C.c_size_t, C.c_size_t, _p, C.c_size_t)
        for nm in ("oct", "quat", "exp"):
            _bind(lib, f"vkgv_meshopt_decode_filter_{nm}", None, _p, C.c_size_t, C.c_size_t)
        _bind(lib, "vkgv_meshopt_decode", C.c_int, _p, C.c_size_t, C.c_size_t, _p, C.c_size_t, C.c_int, C.c_int)
        _host = lib
    return _host


HIP_SYMBOLS = [
    "vkgv_create", "vkgv_destroy", "vkgv_last_error", "vkgv_sync", "vkgv_upload_scene", "vkgv_set_draws",
    "vkgv_set_transforms", "vkgv_view_create", "vkgv_view_destroy", "vkgv_render", "vkgv_view_build_pyramid",
    "vkgv_view_reset_history", "vkgv_view_pyramid_info", "vkgv_view_read_visbuffer", "vkgv_view_read_ids_depth",
    "vkgv_view_read_color", "vkgv_view_read_pyramid", "vkgv_view_write_pyramid", "vkgv_view_write_depth",
    "vkgv_view_read_visible", "vkgv_view_read_stats", "vkgv_set_tuning", "vkgv_get_tuning", "vkgv_srgb_thresholds",
    "vkgv_batch_create", "vkgv_batch_destroy", "vkgv_render_batch", "vkgv_batch_read_stats",
    "vkgv_view_aa_image", "vkgv_view_aa_image_noalloc", "vkgv_view_aa_backward", "vkgv_upload_textures",
    "vkgv_set_hierarchy",
    "vkgv_update_transforms", "vkgv_read_transforms", "vkgv_view_read_motion", "vkgv_measure_atomic_rate",
    "vkgv_upload_scene_device", "vkgv_set_transforms_device", "vkgv_autotune_batch",
]


def _one_hip_runtime():
    """PyTorch-ROCm wheels bundle their own libamdhip64.so.7 and
    libhsa-runtime64.so.  When this library's HIP runtime (the system ROCm)
    is loaded first, a later torch import binds torch to it by soname while
    loading its own HSA runtime, and torch then finds no GPU.  Loading torch
    first (when it is installed) makes the library share torch's runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def hip():
    """libvkgv_hip.so, the HIP frame.  Raises if it is not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is not built: run __graft_entry__.build()")
        _one_hip_runtime()
        lib = C.CDLL(HIP_LIB)
        _bind(lib, "vkgv_create", C.c_int, C.POINTER(CreateInfo), C.POINTER(_p))
        _bind(lib, "vkgv_destroy", None, _p)
        _bind(lib, "vkgv_last_error", C.c_char_p, _p)
        _bind(lib, "vkgv_sync", C.c_int, _p)
        _bind(lib, "vkgv_set_hierarchy", C.c_int, _p, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _p, _u32, _p,
              _u32)
        _bind(lib, "vkgv_update_transforms", C.c_int, _p, C.c_float, _p)
        _bind(lib, "vkgv_read_transforms", C.c_int, _p, _p, _u32)
        _bind(lib, "vkgv_view_read_motion", C.c_int, _p, _p, _p)
        _bind(lib, "vkgv_upload_textures", C.c_int, _p, C.POINTER(TextureDesc), _u32)
        _bind(lib, "vkgv_upload_scene", C.c_int, _p, C.POINTER(MeshDesc), _u32, C.POINTER(PrimitiveDesc), _u32,
              C.POINTER(Material), _u32)
        _bind(lib, "vkgv_set_draws", C.c_int, _p, _p, _u32)
        _bind(lib, "vkgv_set_transforms", C.c_int, _p, _p, _u32)
        _bind(lib, "vkgv_set_transforms_device", C.c_int, _p, _p, _u32, _p)
        _bind(lib, "vkgv_upload_scene_device", C.c_int, _p, C.POINTER(MeshDesc), _u32, C.POINTER(PrimitiveDesc), _u32,
              C.POINTER(Material), _u32)
        _bind(lib, "vkgv_view_create", C.c_int, _p, _u32, _u32, C.POINTER(_p))
        _bind(lib, "vkgv_view_destroy", None, _p)
        _bind(lib, "vkgv_render", C.c_int, _p, _p, C.POINTER(FrameDesc))
        _bind(lib, "vkgv_view_build_pyramid", C.c_int, _p, _p, _p)
        _bind(lib, "vkgv_view_reset_history", C.c_int, _p, _p)
        _bind(lib, "vkgv_view_pyramid_info", C.c_int, _p, C.POINTER(_u32), _p, C.POINTER(_u64))
        for nm in ("read_visbuffer", "read_color", "read_pyramid", "write_pyramid", "write_depth"):
            _bind(lib, f"vkgv_view_{nm}", C.c_int, _p, _p, _p)
        _bind(lib, "vkgv_view_read_ids_depth", C.c_int, _p, _p, _p, _p)
        _bind(lib, "vkgv_view_read_visible", C.c_int, _p, _p, _u32, _p, _u32, C.POINTER(_u32))
        _bind(lib, "vkgv_view_read_stats", C.c_int, _p, _p, C.POINTER(FrameStats))
        _bind(lib, "vkgv_set_tuning", C.c_int, _p, C.POINTER(Tuning))
        _bind(lib, "vkgv_get_tuning", C.c_int, _p, C.POINTER(Tuning))
        _bind(lib, "vkgv_srgb_thresholds", None, _p)
        _bind(lib, "vkgv_measure_atomic_rate", C.c_int, _p, _u32, C.POINTER(C.c_double))
        _bind(lib, "vkgv_batch_create", C.c_int, _p, _p, _u32, C.POINTER(_p))
        _bind(lib, "vkgv_batch_destroy", None, _p)
        _bind(lib, "vkgv_render_batch", C.c_int, _p, _p, _p, _u32)
        _bind(lib, "vkgv_batch_read_stats", C.c_int, _p, _p, C.POINTER(FrameStats))
        _bind(lib, "vkgv_autotune_batch", C.c_int, _p, _p, _p, _u32, _u32, C.POINTER(Tuning), C.POINTER(C.c_float))
        _bind(lib, "vkgv_view_aa_image", C.c_int, _p, _p, C.POINTER(Camera), _p, _p)
        _bind(lib, "vkgv_view_aa_image_noalloc", C.c_int, _p, _p, C.POINTER(Camera), _p, _p)
        _bind(lib, "vkgv_view_aa_backward", C.c_int, _p, _p, C.POINTER(Camera), _p, _p, _u32, _p)
        _hip = lib
    return _hip
