"""The oracle: ctypes bindings of oracle/build/libvkgv_oracle.so (cpuref.c).

Test infrastructure only.  Imported by tests/, by the smoke check in
__graft_entry__.py and by bench.py's CPU-baseline leg, always as the checker
or the CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VKGV_ORACLE_LIB selects another build of the oracle (the sanitized one, make sanitize)
LIB = os.environ.get("VKGV_ORACLE_LIB") or os.path.join(HERE, "build", "libvkgv_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _setup(_lib)
    return _lib


class RefTexture(C.Structure):
    """vkgv_texture_desc."""
    _fields_ = [("rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("wrap_s", C.c_uint32), ("wrap_t", C.c_uint32)]


def _textures(texs):
    keep = [np.ascontiguousarray(t.rgba, dtype=np.uint8) for t in texs]
    arr = (RefTexture * max(1, len(texs)))()
    for i, (t, a) in enumerate(zip(texs, keep)):
        arr[i] = RefTexture(a.ctypes.data, a.shape[1], a.shape[0], int(t.wrap_s), int(t.wrap_t))
    return arr, keep


class RefScene(C.Structure):
    _fields_ = [("meshes", C.c_void_p), ("mesh_count", C.c_uint32),
                ("primitives", C.c_void_p), ("primitive_count", C.c_uint32),
                ("materials", C.c_void_p), ("material_count", C.c_uint32),
                ("draws", C.c_void_p), ("draw_count", C.c_uint32),
                ("transforms", C.c_void_p), ("transform_count", C.c_uint32),
                ("textures", C.c_void_p), ("texture_count", C.c_uint32)]


def _setup(L):
    p = C.c_void_p
    u32 = C.c_uint32
    f32 = C.c_float
    sig = {
        "ref_mat4_mul": (None, [p, p, p]),
        "ref_perspective_rh_zo": (None, [f32, f32, f32, f32, p]),
        "ref_ortho_rh_zo": (None, [f32, f32, f32, f32, f32, f32, p]),
        "ref_look_at_rh": (None, [p, p, p, p]),
        "ref_reverse_depth": (None, [p, p]),
        "ref_frustum_planes": (None, [p, p]),
        "ref_make_camera": (None, [p, p, p]),
        "ref_aabb_in_frustum": (C.c_int, [p, p, p]),
        "ref_world_aabb": (None, [p, p, p, p, p]),
        "ref_project_aabb": (None, [p, p, p, p, p]),
        "ref_pyramid_levels": (u32, [u32, u32, p, p]),
        "ref_pyramid_probe": (f32, [p, u32, u32, p, f32]),
        "ref_floor_log2": (f32, [f32]),
        "ref_occlusion_visible": (C.c_int, [p, p, p, p, u32, u32]),
        "ref_cone_culled": (C.c_int, [p, p, p]),
        "ref_cull": (None, [p, p, p, p, u32, u32, u32, p]),
        "ref_hiz_build": (None, [p, u32, u32, p]),
        "ref_keys_clear": (None, [p, C.c_uint64]),
        "ref_threads": (C.c_int, [C.c_int]),
        "ref_raster_stats": (None, [C.c_int, p, p]),
        "ref_raster": (None, [p, p, p, u32, u32, u32, u32, p]),
        "ref_raster_triangle": (None, [p, p, p, u32, u32, u32, p]),
        "ref_keys_depth": (None, [p, C.c_uint64, p]),
        "ref_keys_ids": (None, [p, C.c_uint64, p]),
        "ref_from_linear": (f32, [f32]),
        "ref_srgb_codes": (None, [p, C.c_uint64, p]),
        "ref_diff_raster": (None, [p, p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, p, p]),
        "ref_diff_forward": (C.c_double, [p, p, C.c_int, p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          p, p, p, C.c_int, p]),
        "ref_diff_backward": (None, [p, p, C.c_int, p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     p, p, p, C.c_int, p, p]),
        "ref_diff_loss_at": (C.c_double, [p, p, C.c_int, p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          p, p, p, C.c_int, p, C.c_int64, p, C.c_int]),
        "ref_frame_aa_image": (None, [p, p, p, C.c_int, C.c_int, p, p]),
        "ref_frame_aa_backward": (None, [p, p, p, C.c_int, C.c_int, p, p, p]),
        "ref_frame_aa_loss_at": (C.c_double, [p, p, p, C.c_int, C.c_int, p, p, p, C.c_int64, p, C.c_int]),
        "ref_resolve": (None, [p, p, p, u32, u32, u32, p]),
        "ref_series_acos": (C.c_double, [C.c_double]),
        "ref_series_sin": (C.c_double, [C.c_double]),
        "ref_anim_sample": (None, [p, p, p, C.c_int, f32, p]),
        "ref_update_transforms": (C.c_int, [p, u32, p, p, u32, p, p, p, p, u32, f32, p, p]),
        "ref_motion_vectors": (None, [p, p, p, p, u32, u32, p]),
        "ref_frame_reference": (u32, [p, p, p, u32, u32, u32, p, p, p, p]),
        "ref_frame_two_pass": (None, [p, p, p, u32, u32, u32, p, p, p, p, p, p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Scene:
    """Oracle view of a SceneArrays (keeps the arrays alive)."""

    def __init__(self, sc):
        self.sc = sc
        self._mesh_arr, self._keep = sc.mesh_descs()
        self._prims = np.ascontiguousarray(sc.primitives)
        self._mats = np.ascontiguousarray(sc.materials)
        self._draws = np.ascontiguousarray(sc.draws)
        self._tr = np.ascontiguousarray(sc.transforms, dtype=np.float32)
        texs = getattr(sc, "textures", [])
        self._tex, self._tex_keep = _textures(texs)
        self.s = RefScene(C.cast(self._mesh_arr, C.c_void_p), len(sc.meshes),
                          self._prims.ctypes.data, len(self._prims), self._mats.ctypes.data, len(self._mats),
                          self._draws.ctypes.data, len(self._draws), self._tr.ctypes.data, self._tr.shape[0],
                          C.cast(self._tex, C.c_void_p), len(texs))

    @property
    def ref(self):
        return C.byref(self.s)

    @property
    def draw_count(self):
        return len(self._draws)


def threads(n: int = 0) -> int:
    """OpenMP threads of the frame functions; n > 0 sets the count."""
    return int(lib().ref_threads(int(n)))


def raster_stats(reset: bool = True):
    """(triangles past the back-face test, fragments) rasterized since the
    last reset; reset zeroes the counters."""
    t, f = C.c_uint64(), C.c_uint64()
    lib().ref_raster_stats(1 if reset else 0, C.byref(t), C.byref(f))
    return t.value, f.value


def pyramid_info(w, h):
    sizes = np.zeros(64, dtype=np.uint32)
    offs = np.zeros(33, dtype=np.uint64)
    n = lib().ref_pyramid_levels(w, h, sizes.ctypes.data, offs.ctypes.data)
    return n, [(int(sizes[2 * k]), int(sizes[2 * k + 1])) for k in range(n)], int(offs[n])


def hiz_build(depth: np.ndarray) -> np.ndarray:
    h, w = depth.shape
    _, _, total = pyramid_info(w, h)
    out = np.zeros(total, dtype=np.float32)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    lib().ref_hiz_build(d.ctypes.data, w, h, out.ctypes.data)
    return out


def cull(scene: Scene, camera, pyramid, w, h, flags, eye=None) -> np.ndarray:
    vis = np.zeros(scene.draw_count, dtype=np.uint8)
    pyr = np.ascontiguousarray(pyramid, dtype=np.float32)
    e = np.asarray(eye if eye is not None else (0, 0, 0), dtype=np.float32)
    lib().ref_cull(scene.ref, C.byref(camera), e.ctypes.data, pyr.ctypes.data, w, h, flags, vis.ctypes.data)
    return np.nonzero(vis)[0].astype(np.uint32)


def raster(scene: Scene, vp, draw_list, w, h, flags=0, keys=None) -> np.ndarray:
    if keys is None:
        keys = np.full(w * h, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    lst = np.ascontiguousarray(draw_list, dtype=np.uint32)
    vpa = np.ascontiguousarray(np.asarray(vp, dtype=np.float32).ravel())
    lib().ref_raster(scene.ref, vpa.ctypes.data, lst.ctypes.data, lst.size, w, h, flags, keys.ctypes.data)
    return keys.reshape(h, w)


def resolve(scene: Scene, vp, keys, flags=0) -> np.ndarray:
    h, w = keys.shape
    out = np.zeros(w * h, dtype=np.uint32)
    vpa = np.ascontiguousarray(np.asarray(vp, dtype=np.float32).ravel())
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    lib().ref_resolve(scene.ref, vpa.ctypes.data, k.ctypes.data, w, h, flags, out.ctypes.data)
    return out.reshape(h, w)


def srgb_codes(c) -> np.ndarray:
    c = np.ascontiguousarray(c, dtype=np.float32)
    out = np.zeros(c.size, dtype=np.uint8)
    lib().ref_srgb_codes(c.ctypes.data, c.size, out.ctypes.data)
    return out


def raster_triangle(c0, c1, c2, tid, w, h, keys=None) -> np.ndarray:
    """One clip-space triangle (x, y, z, w per vertex) into `keys` (h, w)."""
    if keys is None:
        keys = np.full(w * h, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    v = [np.ascontiguousarray(c, dtype=np.float32) for c in (c0, c1, c2)]
    lib().ref_raster_triangle(v[0].ctypes.data, v[1].ctypes.data, v[2].ctypes.data, tid, w, h, keys.ctypes.data)
    return keys.reshape(h, w)


def keys_depth(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    return (~(k >> np.uint64(32)).astype(np.uint32)).view(np.float32)


def keys_ids(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    return (np.uint32(0xFFFFFFFE) - (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)).astype(np.uint32)


def frame_reference(scene: Scene, camera, w, h, flags, pyramid, eye=None):
    """Reference frame; `pyramid` (float array) is read then overwritten."""
    keys = np.empty(w * h, dtype=np.uint64)
    rgba = np.zeros(w * h, dtype=np.uint32)
    vis = np.zeros(scene.draw_count, dtype=np.uint8)
    e = np.asarray(eye if eye is not None else (0, 0, 0), dtype=np.float32)
    lib().ref_frame_reference(scene.ref, C.byref(camera), e.ctypes.data, w, h, flags, pyramid.ctypes.data,
                              keys.ctypes.data, rgba.ctypes.data, vis.ctypes.data)
    return keys.reshape(h, w), rgba.reshape(h, w), np.nonzero(vis)[0].astype(np.uint32)


def frame_two_pass(scene: Scene, camera, w, h, flags, history, eye=None):
    """Two-pass frame; `history` (u8 per draw) is read and updated."""
    _, _, total = pyramid_info(w, h)
    pyr = np.zeros(total, dtype=np.float32)
    keys = np.empty(w * h, dtype=np.uint64)
    rgba = np.zeros(w * h, dtype=np.uint32)
    ve = np.zeros(scene.draw_count, dtype=np.uint8)
    vl = np.zeros(scene.draw_count, dtype=np.uint8)
    e = np.asarray(eye if eye is not None else (0, 0, 0), dtype=np.float32)
    lib().ref_frame_two_pass(scene.ref, C.byref(camera), e.ctypes.data, w, h, flags, history.ctypes.data,
                             pyr.ctypes.data, keys.ctypes.data, rgba.ctypes.data, ve.ctypes.data, vl.ctypes.data)
    return (keys.reshape(h, w), rgba.reshape(h, w), pyr,
            np.nonzero(ve)[0].astype(np.uint32), np.nonzero(vl)[0].astype(np.uint32))


# ---------------------------------------------------------------------------
# Differentiable form (diffref.c)

def diff_raster(clip, tris, H, W):
    """clip [C, I, V, 4] float32, tris [T, 3] int32 -> (ids [C, H, W] int32, depth float32)."""
    c = np.ascontiguousarray(clip, dtype=np.float32)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    C_, I_, V_ = c.shape[:3]
    ids = np.zeros((C_, H, W), np.int32)
    depth = np.zeros((C_, H, W), np.float32)
    lib().ref_diff_raster(c.ctypes.data, t.ctypes.data, C_, I_, V_, t.shape[0], H, W, ids.ctypes.data, depth.ctypes.data)
    return ids, depth


def _diff_args(clip, attr, tris, ids, depth, target):
    c = np.ascontiguousarray(clip, dtype=np.float64)
    a = np.ascontiguousarray(attr, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.int32)
    i = np.ascontiguousarray(ids, dtype=np.int32)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    g = None if target is None else np.ascontiguousarray(target, dtype=np.float64)
    C_, I_, V_ = c.shape[:3]
    H, W = i.shape[1:]
    return (c, a, t, i, d, g), (c.ctypes.data, a.ctypes.data, a.shape[1], t.ctypes.data, C_, I_, V_, t.shape[0], H, W,
                                 i.ctypes.data, d.ctypes.data, None if g is None else g.ctypes.data)


def diff_forward(clip, attr, tris, ids, depth, target=None, aa=True):
    """(loss, image [C, H, W, A]) in float64 for fixed ids / depth."""
    keep, args = _diff_args(clip, attr, tris, ids, depth, target)
    C_ = keep[0].shape[0]
    H, W = keep[3].shape[1:]
    out = np.zeros((C_, H, W, keep[1].shape[1]), np.float64)
    loss = lib().ref_diff_forward(*args, 1 if aa else 0, out.ctypes.data)
    return loss, out


def diff_backward(clip, attr, tris, ids, depth, target=None, aa=True):
    """(d loss / d clip [C, I, V, 4], d loss / d attr [V, A]) in float64."""
    keep, args = _diff_args(clip, attr, tris, ids, depth, target)
    g_clip = np.zeros(keep[0].shape, np.float64)
    g_attr = np.zeros(keep[1].shape, np.float64)
    lib().ref_diff_backward(*args, 1 if aa else 0, g_clip.ctypes.data, g_attr.ctypes.data)
    return g_clip, g_attr


def diff_loss_at(clip, attr, tris, ids, depth, pix, target=None, aa=True, state=None, mode=0):
    """Loss over the listed pixels (flat indices into [C, H, W]).  state:
    int8 [len(pix), 4] antialiasing state, recorded (mode 1) or held fixed
    (mode 2)."""
    keep, args = _diff_args(clip, attr, tris, ids, depth, target)
    px = np.ascontiguousarray(pix, dtype=np.int64)
    st = None
    if state is not None:
        assert state.dtype == np.int8 and state.shape == (px.size, 4) and state.flags.c_contiguous
        st = state.ctypes.data
    return lib().ref_diff_loss_at(*args, 1 if aa else 0, px.ctypes.data, px.size, st, mode)


def _frame_args(keys, vp, xf):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(vp, dtype=np.float32).ravel()
    x = None if xf is None else np.ascontiguousarray(xf, dtype=np.float64)
    return k, v, x


def frame_aa_image(scene: Scene, vp, keys, W, H, xf=None):
    """Antialiased linear image [H, W, 3] (float64) of a frame's keys; vp is
    the camera's column-major view-projection, xf optional float64
    transforms [T, 16] (default: the scene's)."""
    k, v, x = _frame_args(keys, vp, xf)
    out = np.zeros((H, W, 3), np.float64)
    lib().ref_frame_aa_image(scene.ref, v.ctypes.data, k.ctypes.data, W, H, _ptr(x), out.ctypes.data)
    return out


def frame_aa_backward(scene: Scene, vp, keys, W, H, g_image, xf=None):
    """d(sum g_image * image) / d transforms, float64 [T, 16] (column-major)."""
    k, v, x = _frame_args(keys, vp, xf)
    g = np.ascontiguousarray(g_image, dtype=np.float64)
    out = np.zeros((scene.sc.transforms.reshape(-1, 16).shape[0], 16), np.float64)
    lib().ref_frame_aa_backward(scene.ref, v.ctypes.data, k.ctypes.data, W, H, _ptr(x), g.ctypes.data, out.ctypes.data)
    return out


def frame_aa_loss_at(scene: Scene, vp, keys, W, H, xf, target, pix, state=None, mode=0):
    """sum over the listed pixels of |image - target|^2; state int8 [len(pix), 4]."""
    k, v, x = _frame_args(keys, vp, xf)
    t = np.ascontiguousarray(target, dtype=np.float64)
    px = np.ascontiguousarray(pix, dtype=np.int64)
    st = None
    if state is not None:
        assert state.dtype == np.int8 and state.shape == (px.size, 4) and state.flags.c_contiguous
        st = state.ctypes.data
    return lib().ref_frame_aa_loss_at(scene.ref, v.ctypes.data, k.ctypes.data, W, H, _ptr(x), t.ctypes.data,
                                      px.ctypes.data, px.size, st, mode)


def series_acos(x):
    return lib().ref_series_acos(float(x))


def series_sin(a):
    return lib().ref_series_sin(float(a))


def anim_sample(sampler, times, values, n, time):
    """AnimationSampler::sample of one sampler (a 1-element ANIM_SAMPLER_DTYPE record)."""
    sp = np.ascontiguousarray(sampler)
    t = np.ascontiguousarray(times, np.float32)
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(4, np.float32)
    lib().ref_anim_sample(sp.ctypes.data, t.ctypes.data, v.ctypes.data, n, float(time), out.ctypes.data)
    return out[:n]


def update_transforms(h, time):
    """World matrices of the mesh nodes of hierarchy `h` at `time`, (T, 16) column-major."""
    nodes = np.ascontiguousarray(h.nodes)
    ch = np.ascontiguousarray(h.children, np.uint32)
    rt = np.ascontiguousarray(h.roots, np.uint32)
    sp = np.ascontiguousarray(h.samplers)
    tm = np.ascontiguousarray(h.times, np.float32)
    vl = np.ascontiguousarray(h.values, np.float32)
    cn = np.ascontiguousarray(h.channels)
    cnt = C.c_uint32(0)
    ptr = lambda a: a.ctypes.data if a.size else None
    args = [ptr(nodes), len(nodes), ptr(ch), ptr(rt), len(rt), ptr(sp), ptr(tm), ptr(vl), ptr(cn), len(cn), float(time)]
    if lib().ref_update_transforms(*args, None, C.byref(cnt)) != 0:
        raise ValueError("hierarchy walk failed")
    out = np.zeros((cnt.value, 16), np.float32)
    lib().ref_update_transforms(*args, out.ctypes.data if out.size else None, C.byref(cnt))
    return out


def motion_vectors(scene: Scene, camera, prev_transforms, keys, w, h):
    """Motion vectors (H, W, 2) of a frame's keys."""
    k = np.ascontiguousarray(keys, np.uint64).ravel()
    pt = np.ascontiguousarray(prev_transforms, np.float32)
    out = np.zeros(w * h * 2, np.float32)
    lib().ref_motion_vectors(scene.ref, C.byref(camera), pt.ctypes.data, k.ctypes.data, w, h, out.ctypes.data)
    return out.reshape(h, w, 2)
