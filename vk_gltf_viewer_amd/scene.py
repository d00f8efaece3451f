"""Host-side scene, camera and meshlet producers (libvkgv_host.so).

`Scene` owns a generated scene and exposes its contract arrays as numpy views
(valid while the Scene lives).  `Scene.arrays()` returns copies that outlive
it.  The frame's inputs are exactly these arrays: meshes (buffer sets),
primitives, materials, draws (World::rebuildDrawBuffer, world.cpp:230-293),
transforms (World::updateTransformBuffer, world.cpp:295-345) and views.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

MESHLET_DTYPE = np.dtype({"names": ["vertex_offset", "triangle_offset", "vertex_count", "triangle_count", "pad",
                                    "aabb_extents", "aabb_center"],
                          "formats": ["<u4", "<u4", "u1", "u1", ("u1", 2), ("<f4", 3), ("<f4", 3)],
                          "offsets": [0, 4, 8, 9, 10, 12, 24], "itemsize": 36})
VERTEX_DTYPE = np.dtype({"names": ["position", "color", "normal", "pad", "uv"],
                         "formats": [("<f4", 3), ("u1", 4), ("u1", 3), "u1", ("<u2", 2)],
                         "offsets": [0, 12, 16, 19, 20], "itemsize": 24})
DRAW_DTYPE = np.dtype({"names": ["primitive_index", "meshlet_index", "transform_index"],
                       "formats": ["<u4", "<u4", "<u4"], "offsets": [0, 4, 8], "itemsize": 12})
MATERIAL_DTYPE = np.dtype({"names": ["albedo_factor", "albedo_index", "uv_offset", "uv_scale", "uv_rotation",
                                     "alpha_cutoff", "double_sided"],
                           "formats": [("<f4", 4), "<u4", ("<f4", 2), ("<f4", 2), "<f4", "<f4", "<u4"],
                           "offsets": [0, 16, 20, 28, 36, 40, 44], "itemsize": 48})
CONE_DTYPE = np.dtype({"names": ["apex", "axis", "cutoff", "pad"],
                       "formats": [("<f4", 3), ("<f4", 3), "<f4", "<f4"],
                       "offsets": [0, 12, 24, 28], "itemsize": 32})
PRIM_DTYPE = np.dtype({"names": ["mesh_index", "material_index", "aabb_extents", "aabb_center"],
                       "formats": ["<u4", "<u4", ("<f4", 3), ("<f4", 3)],
                       "offsets": [0, 4, 8, 20], "itemsize": 32})


NODE_DTYPE = np.dtype({"names": ["mesh", "first_child", "child_count", "pad", "translation", "rotation", "scale"],
                       "formats": ["<i4", "<u4", "<u4", "<u4", ("<f4", 3), ("<f4", 4), ("<f4", 3)],
                       "offsets": [0, 4, 8, 12, 16, 28, 44], "itemsize": 56})
ANIM_SAMPLER_DTYPE = np.dtype({"names": ["input_offset", "input_count", "output_offset", "interpolation"],
                               "formats": ["<u4"] * 4, "offsets": [0, 4, 8, 12], "itemsize": 16})
ANIM_CHANNEL_DTYPE = np.dtype({"names": ["node", "sampler", "path", "pad"],
                               "formats": ["<u4"] * 4, "offsets": [0, 4, 8, 12], "itemsize": 16})


def _view(ptr, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_uint8 * (count * dtype.itemsize)).from_address(C.cast(ptr, C.c_void_p).value)
    return np.frombuffer(buf, dtype=dtype, count=count)


@dataclass
class MeshArrays:
    vertex_indices: np.ndarray      # u32
    triangle_indices: np.ndarray    # u8, 3 per triangle, 4-byte padded per meshlet
    vertices: np.ndarray            # VERTEX_DTYPE
    meshlets: np.ndarray            # MESHLET_DTYPE
    cones: np.ndarray | None = None  # CONE_DTYPE


@dataclass
class Texture:
    """An albedo texture: sRGB-encoded RGBA8, row 0 at v = 0, and the glTF
    sampler's wrap modes (WRAP_REPEAT / WRAP_CLAMP_TO_EDGE / WRAP_MIRRORED_REPEAT)."""
    rgba: np.ndarray                # (H, W, 4) uint8
    wrap_s: int = 0
    wrap_t: int = 0


def texture_descs(textures):
    """ctypes vkgv_texture_desc array of a texture list, and the arrays it
    points into (keep them alive while the descriptors are used)."""
    keep = [np.ascontiguousarray(t.rgba, dtype=np.uint8) for t in textures]
    arr = (N.TextureDesc * max(1, len(textures)))()
    for i, (t, a) in enumerate(zip(textures, keep)):
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError(f"texture {i}: expected (H, W, 4) uint8")
        arr[i] = N.TextureDesc(a.ctypes.data, a.shape[1], a.shape[0], int(t.wrap_s), int(t.wrap_t))
    return arr, keep


@dataclass
class Hierarchy:
    """glTF nodes of a scene and their animations, for the device transform
    update (vkgv_set_hierarchy / vkgv_update_transforms)."""
    nodes: np.ndarray                                   # NODE_DTYPE
    children: np.ndarray                                # u32
    roots: np.ndarray                                   # u32
    samplers: np.ndarray = field(default_factory=lambda: np.zeros(0, ANIM_SAMPLER_DTYPE))
    times: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    values: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    channels: np.ndarray = field(default_factory=lambda: np.zeros(0, ANIM_CHANNEL_DTYPE))

    @property
    def duration(self) -> float:
        """Longest sampler period (the last key time)."""
        if not len(self.samplers):
            return 0.0
        last = self.samplers["input_offset"].astype(np.int64) + self.samplers["input_count"] - 1
        return float(self.times[last].max())


@dataclass
class SceneArrays:
    meshes: list[MeshArrays]
    primitives: np.ndarray          # PRIM_DTYPE
    materials: np.ndarray           # MATERIAL_DTYPE
    draws: np.ndarray               # DRAW_DTYPE
    transforms: np.ndarray          # (T, 16) f32, column-major
    views: list = field(default_factory=list)
    textures: list = field(default_factory=list)   # Texture, indexed by the materials' albedo_index
    hierarchy: Hierarchy | None = None              # node tree + animations (glTF scenes)

    @property
    def triangle_count(self) -> int:
        n = 0
        for m in self.meshes:
            pass
        tc = [m.meshlets["triangle_count"].astype(np.int64) for m in self.meshes]
        pm = self.primitives["mesh_index"]
        d = self.draws
        for mi, counts in enumerate(tc):
            sel = pm[d["primitive_index"]] == mi
            n += int(counts[d["meshlet_index"][sel]].sum())
        return n

    def mesh_descs(self):
        """(MeshDesc array, keep-alive list) for the C ABI."""
        arr = (N.MeshDesc * max(1, len(self.meshes)))()
        keep = []
        for i, m in enumerate(self.meshes):
            parts = [np.ascontiguousarray(m.vertex_indices, dtype=np.uint32),
                     np.ascontiguousarray(m.triangle_indices, dtype=np.uint8),
                     np.ascontiguousarray(m.vertices), np.ascontiguousarray(m.meshlets)]
            cones = np.ascontiguousarray(m.cones) if m.cones is not None else None
            keep.extend(parts)
            if cones is not None:
                keep.append(cones)
            d = arr[i]
            d.vertex_indices = parts[0].ctypes.data
            d.vertex_index_count = parts[0].size
            d.triangle_indices = parts[1].ctypes.data
            d.triangle_index_bytes = parts[1].size
            d.vertices = parts[2].ctypes.data
            d.vertex_count = parts[2].size
            d.meshlets = parts[3].ctypes.data
            d.meshlet_count = parts[3].size
            d.cones = cones.ctypes.data if cones is not None else None
        return arr, keep


class Scene:
    """A seeded synthetic scene generated by libvkgv_host.so."""

    def __init__(self, config: int, seed: int, variant: int = 0):
        lib = N.host()
        self._lib = lib
        self._h = lib.vkgv_scene_generate(config, seed, variant)
        if not self._h:
            raise RuntimeError("scene generation failed: " + lib.vkgv_scene_error().decode())
        self.config, self.seed, self.variant = config, seed, variant

    def close(self):
        if self._h:
            self._lib.vkgv_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self) -> SceneArrays:
        lib, h = self._lib, self._h
        meshes = []
        md = lib.vkgv_scene_meshes(h)
        for i in range(lib.vkgv_scene_mesh_count(h)):
            d = md[i]
            meshes.append(MeshArrays(
                vertex_indices=_view(d.vertex_indices, d.vertex_index_count, np.dtype("<u4")).copy(),
                triangle_indices=_view(d.triangle_indices, d.triangle_index_bytes, np.dtype("u1")).copy(),
                vertices=_view(d.vertices, d.vertex_count, VERTEX_DTYPE).copy(),
                meshlets=_view(d.meshlets, d.meshlet_count, MESHLET_DTYPE).copy(),
                cones=_view(d.cones, d.meshlet_count, CONE_DTYPE).copy() if d.cones else None))
        prims = _view(lib.vkgv_scene_primitives(h), lib.vkgv_scene_primitive_count(h), PRIM_DTYPE).copy()
        mats = _view(lib.vkgv_scene_materials(h), lib.vkgv_scene_material_count(h), MATERIAL_DTYPE).copy()
        draws = _view(lib.vkgv_scene_draws(h), lib.vkgv_scene_draw_count(h), DRAW_DTYPE).copy()
        nt = lib.vkgv_scene_transform_count(h)
        tr = _view(lib.vkgv_scene_transforms(h), nt * 16, np.dtype("<f4")).copy().reshape(nt, 16)
        vp = lib.vkgv_scene_views(h)
        views = [N.ViewDesc.from_buffer_copy(vp[i]) for i in range(lib.vkgv_scene_view_count(h))]
        return SceneArrays(meshes, prims, mats, draws, tr, views)

    @property
    def triangle_count(self) -> int:
        return int(self._lib.vkgv_scene_triangle_count(self._h))


def camera_from_view(view: N.ViewDesc) -> N.Camera:
    """Camera buffer of a first frame (camera.cpp:170-193)."""
    cam = N.Camera()
    N.host().vkgv_camera_from_view(C.byref(view), C.byref(cam))
    return cam


def camera_advance(prev: N.Camera, view: N.ViewDesc) -> N.Camera:
    """Next frame's camera: prev* <- current, new matrices (camera.cpp:181-193)."""
    cam = N.Camera()
    N.host().vkgv_camera_advance(C.byref(prev), C.byref(view), C.byref(cam))
    return cam


def build_meshlets(positions: np.ndarray, indices: np.ndarray, max_vertices: int = 64,
                   max_triangles: int = 124) -> MeshArrays:
    """Meshlets at the reference's limits (assets.cpp:322-373); positions (V,3) f32."""
    lib = N.host()
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    out = N.MeshletBuild()
    rc = lib.vkgv_build_meshlets(pos.ctypes.data, pos.shape[0], idx.ctypes.data, idx.size, max_vertices,
                                 max_triangles, C.byref(out))
    if rc != 0:
        raise ValueError("meshlet build failed: " + lib.vkgv_scene_error().decode())
    try:
        verts = np.zeros(pos.shape[0], dtype=VERTEX_DTYPE)
        verts["position"] = pos
        verts["color"] = 255
        res = MeshArrays(
            vertex_indices=np.ctypeslib.as_array(out.vertex_indices, (out.vertex_index_count,)).copy(),
            triangle_indices=np.ctypeslib.as_array(out.triangle_indices, (out.triangle_index_bytes,)).copy(),
            vertices=verts,
            meshlets=_view(out.meshlets, out.meshlet_count, MESHLET_DTYPE).copy(),
            cones=_view(out.cones, out.meshlet_count, CONE_DTYPE).copy())
    finally:
        lib.vkgv_meshlet_build_free(C.byref(out))
    return res


def view_desc(width, height, eye, target, fovy_deg=60.0, up=(0.0, 1.0, 0.0), znear=0.1, zfar=1000.0,
              ortho_half=None) -> N.ViewDesc:
    v = N.ViewDesc()
    v.width, v.height = width, height
    v.eye[:] = list(eye)
    v.target[:] = list(target)
    v.up[:] = list(up)
    v.fovy = np.float32(np.deg2rad(fovy_deg))
    v.znear, v.zfar = znear, zfar
    if ortho_half is not None:
        v.orthographic = 1
        v.ortho_half_width, v.ortho_half_height = ortho_half
    return v
