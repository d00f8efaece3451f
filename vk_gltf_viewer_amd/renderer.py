"""Host-side mirror of the reference's frame interface over libvkgv_hip.so.

The reference drives its frame from ``Application::run()`` with a World
(draws, transforms, primitives, materials) and a Camera
(application.cpp:604-1280).  Here:

    ctx  = Renderer()                         # Device + World GPU buffers
    ctx.upload_scene(scene_arrays)            # assets.cpp:375-429, world.cpp:10-185
    ctx.set_draws(draws); ctx.set_transforms(mats)   # world.cpp:230-345
    view = ctx.create_view(W, H)              # the pass images (application.cpp:181-576)
    view.render(camera, flags)                # the recorded frame (application.cpp:843-1003)
    view.ids_depth(), view.color(), view.pyramid(), view.visible()

Every call goes through the C ABI of include/vkgv.h; there is no Python or CPU
fallback: a missing library or a failing call raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .scene import SceneArrays, texture_descs


class VkgvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vkgv error {code}: {msg}")
        self.code = code


class Renderer:
    def __init__(self, device: int = 0, big_triangle_capacity: int = 0, clip_triangle_capacity: int = 0,
                 tile_triangle_capacity: int = 0):
        """Queue capacities are per view and raster pass (0 = the library's
        defaults; the tile capacity is per 32x32 screen tile); a triangle that
        finds its queue full is rasterized by the slower direct path, so they
        change speed, never the image."""
        self._lib = N.hip()
        self.device = device
        info = N.CreateInfo(device, big_triangle_capacity, clip_triangle_capacity, tile_triangle_capacity)
        h = C.c_void_p()
        rc = self._lib.vkgv_create(C.byref(info), C.byref(h))
        if rc != 0:
            raise VkgvError(rc, "vkgv_create failed (no HIP device?)")
        self._h = h
        self.views = []
        self.batches = []

    # -- errors -----------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise VkgvError(rc, self._lib.vkgv_last_error(self._h).decode())
        return rc

    def close(self):
        for b in list(self.batches):
            b.close()
        for v in list(self.views):
            v.close()
        if self._h:
            self._lib.vkgv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scene ------------------------------------------------------------------
    def upload_scene(self, sc: SceneArrays):
        meshes, keep = sc.mesh_descs()
        prims = np.ascontiguousarray(sc.primitives)
        mats = np.ascontiguousarray(sc.materials)
        self._check(self._lib.vkgv_upload_scene(
            self._h, meshes, len(sc.meshes),
            C.cast(prims.ctypes.data, C.POINTER(N.PrimitiveDesc)), len(prims),
            C.cast(mats.ctypes.data, C.POINTER(N.Material)), len(mats)))
        del keep
        self.upload_textures(sc.textures)

    def upload_scene_device(self, sc: SceneArrays, corrupt=None):
        """vkgv_upload_scene_device: each mesh's arrays are first placed in
        device memory (torch tensors, standing in for a caller's own device
        buffers) and handed over by address.  `corrupt(mesh_index, parts)` may
        modify the device copies (tests of the device-side check)."""
        import torch
        arr = (N.MeshDesc * max(1, len(sc.meshes)))()
        keep = []
        for i, m in enumerate(sc.meshes):
            host = [np.ascontiguousarray(m.vertex_indices, dtype=np.uint32), np.ascontiguousarray(m.triangle_indices,
                                                                                                 dtype=np.uint8),
                    np.ascontiguousarray(m.vertices), np.ascontiguousarray(m.meshlets)]
            if m.cones is not None:
                host.append(np.ascontiguousarray(m.cones))
            dev = [torch.from_numpy(h.view(np.uint8).reshape(-1).copy()).to("cuda") for h in host]
            if corrupt is not None:
                corrupt(i, dev)
            keep.extend(dev)
            d = arr[i]
            d.vertex_indices, d.vertex_index_count = dev[0].data_ptr(), host[0].size
            d.triangle_indices, d.triangle_index_bytes = dev[1].data_ptr(), host[1].size
            d.vertices, d.vertex_count = dev[2].data_ptr(), host[2].size
            d.meshlets, d.meshlet_count = dev[3].data_ptr(), host[3].size
            d.cones = dev[4].data_ptr() if m.cones is not None else None
        torch.cuda.synchronize()          # the copies above ran on torch's stream
        prims = np.ascontiguousarray(sc.primitives)
        mats = np.ascontiguousarray(sc.materials)
        self._check(self._lib.vkgv_upload_scene_device(
            self._h, arr, len(sc.meshes), C.cast(prims.ctypes.data, C.POINTER(N.PrimitiveDesc)), len(prims),
            C.cast(mats.ctypes.data, C.POINTER(N.Material)), len(mats)))
        del keep
        self.upload_textures(sc.textures)

    def set_transforms_device(self, transforms, stream=None):
        """vkgv_set_transforms_device from a contiguous float32 [T, 16] torch
        device tensor, ordered on torch's current stream (fenced when that is
        the default stream, which the library reads as its own)."""
        import torch
        t = transforms
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                and t.dim() == 2 and t.shape[1] == 16):
            raise ValueError("transforms must be a contiguous float32 [T, 16] device tensor")
        s = torch.cuda.current_stream()
        if s.cuda_stream:
            self._check(self._lib.vkgv_set_transforms_device(self._h, t.data_ptr(), t.shape[0], s.cuda_stream))
            return
        s.synchronize()
        self._check(self._lib.vkgv_set_transforms_device(self._h, t.data_ptr(), t.shape[0], None))
        self.sync()

    def upload_textures(self, textures):
        """Albedo textures (scene.Texture list) for VKGV_RESOLVE_TEXTURED."""
        descs, keep = texture_descs(textures)
        self._check(self._lib.vkgv_upload_textures(self._h, descs, len(textures)))
        del keep

    def set_draws(self, draws: np.ndarray):
        d = np.ascontiguousarray(draws)
        self._check(self._lib.vkgv_set_draws(self._h, d.ctypes.data, len(d)))

    def set_transforms(self, transforms: np.ndarray):
        t = np.ascontiguousarray(transforms, dtype=np.float32).reshape(-1, 16)
        self._check(self._lib.vkgv_set_transforms(self._h, t.ctypes.data, t.shape[0]))

    def load(self, sc: SceneArrays):
        self.upload_scene(sc)
        self.set_transforms(sc.transforms)
        self.set_draws(sc.draws)

    def set_hierarchy(self, h):
        """Node tree and animations (scene.Hierarchy) for update_transforms; the
        context's transforms become the mesh nodes' matrices at time 0."""
        nodes, samplers, channels = (np.ascontiguousarray(a) for a in (h.nodes, h.samplers, h.channels))
        children, roots = (np.ascontiguousarray(a, dtype=np.uint32) for a in (h.children, h.roots))
        times, values = (np.ascontiguousarray(a, dtype=np.float32) for a in (h.times, h.values))
        ptr = lambda a: a.ctypes.data if a.size else None
        self._check(self._lib.vkgv_set_hierarchy(
            self._h, ptr(nodes), len(nodes), ptr(children), len(children), ptr(roots), len(roots), ptr(samplers),
            len(samplers),
            ptr(times), len(times), ptr(values), len(values), ptr(channels),
            len(channels)))

    def update_transforms(self, time: float, stream=None):
        """Animation sampling and world matrices on the device (stream-ordered)."""
        self._check(self._lib.vkgv_update_transforms(self._h, float(time), stream))

    def read_transforms(self, count: int) -> np.ndarray:
        out = np.empty((count, 16), dtype=np.float32)
        self._check(self._lib.vkgv_read_transforms(self._h, out.ctypes.data, count))
        return out

    def sync(self):
        self._check(self._lib.vkgv_sync(self._h))

    def set_tuning(self, small_max_extent: int, raster_blocks_per_cu: int, tile_item_tris: int = 8,
                   resolve_cache_slots: int = 8192):
        t = N.Tuning(small_max_extent, raster_blocks_per_cu, tile_item_tris, resolve_cache_slots)
        self._check(self._lib.vkgv_set_tuning(self._h, C.byref(t)))

    def get_tuning(self):
        """(small_max_extent, raster_blocks_per_cu, tile_item_tris, resolve_cache_slots)"""
        t = N.Tuning()
        self._check(self._lib.vkgv_get_tuning(self._h, C.byref(t)))
        return t.small_max_extent, t.raster_blocks_per_cu, t.tile_item_tris, t.resolve_cache_slots

    def measure_atomic_rate(self, scattered: bool = False) -> float:
        """64-bit atomicMin per second on a 4K key buffer (vkgv_measure_atomic_rate)."""
        r = C.c_double()
        self._check(self._lib.vkgv_measure_atomic_rate(self._h, 1 if scattered else 0, C.byref(r)))
        return r.value

    def device_identity(self) -> str:
        """Architecture and compute-unit count of the context's device (the
        autotuner's cache key)."""
        try:
            import torch
            p = torch.cuda.get_device_properties(self.device)
            return f"{p.gcnArchName}-{p.multi_processor_count}cu"
        except Exception:
            return f"hip-device-{self.device}"

    def create_batch(self, views) -> "Batch":
        return Batch(self, views)

    def create_view(self, width: int, height: int) -> "View":
        h = C.c_void_p()
        self._check(self._lib.vkgv_view_create(self._h, width, height, C.byref(h)))
        v = View(self, h, width, height)
        self.views.append(v)
        return v


class Batch:
    """Views rendered together by one sequence of launches (vkgv_render_batch)."""

    def __init__(self, r: Renderer, views):
        self.r, self.views = r, list(views)
        arr = (C.c_void_p * len(self.views))(*[v._h.value for v in self.views])
        h = C.c_void_p()
        r._check(r._lib.vkgv_batch_create(r._h, arr, len(self.views), C.byref(h)))
        self._h = h
        self._frames = (N.FrameDesc * len(self.views))()
        r.batches.append(self)

    def close(self):
        if self._h:
            self.r._lib.vkgv_batch_destroy(self._h)
            self._h = None
            self.r.batches.remove(self)

    def render(self, cameras, flags: int = N.MODE_REFERENCE, eyes=None, stream=None):
        if len(cameras) != len(self.views):
            raise ValueError("one camera per view")
        for i, cam in enumerate(cameras):
            f = self._frames[i]
            f.camera = cam
            f.eye[:] = list(eyes[i]) if eyes is not None else [0.0, 0.0, 0.0]
            f.flags = flags
            f.stream = stream
        self.r._check(self.r._lib.vkgv_render_batch(self.r._h, self._h, self._frames, len(self.views)))

    def stats(self) -> N.FrameStats:
        s = N.FrameStats()
        self.r._check(self.r._lib.vkgv_batch_read_stats(self.r._h, self._h, C.byref(s)))
        return s

    def autotune(self, cameras, flags: int, frames_per_candidate: int = 3, eyes=None):
        """The library's own autotuner (vkgv_autotune_batch) on this batch
        frame: applies and returns the fastest tuning and its ms per frame."""
        for i, cam in enumerate(cameras):
            f = self._frames[i]
            f.camera = cam
            f.eye[:] = list(eyes[i]) if eyes is not None else [0.0, 0.0, 0.0]
            f.flags = flags
            f.stream = None
        t, ms = N.Tuning(), C.c_float()
        self.r._check(self.r._lib.vkgv_autotune_batch(self.r._h, self._h, self._frames, len(self.views),
                                                      frames_per_candidate, C.byref(t), C.byref(ms)))
        return (t.small_max_extent, t.raster_blocks_per_cu, t.tile_item_tris, t.resolve_cache_slots), ms.value


class View:
    def __init__(self, r: Renderer, h, width, height):
        self.r, self._h, self.width, self.height = r, h, width, height
        lv = C.c_uint32()
        n = C.c_uint64()
        sizes = (C.c_uint32 * 64)()
        r._lib.vkgv_view_pyramid_info(h, C.byref(lv), sizes, C.byref(n))
        self.levels = lv.value
        self.level_sizes = [(sizes[2 * k], sizes[2 * k + 1]) for k in range(self.levels)]
        self.pyramid_texels = n.value

    def close(self):
        if self._h:
            self.r._lib.vkgv_view_destroy(self._h)
            self._h = None
            self.r.views.remove(self)

    def render(self, camera: N.Camera, flags: int = N.MODE_REFERENCE, eye=(0.0, 0.0, 0.0), stream=None):
        f = N.FrameDesc()
        f.camera = camera
        f.eye[:] = list(eye)
        f.flags = flags
        f.stream = stream
        self.r._check(self.r._lib.vkgv_render(self.r._h, self._h, C.byref(f)))

    def build_pyramid(self):
        self.r._check(self.r._lib.vkgv_view_build_pyramid(self.r._h, self._h, None))

    # -- differentiable frame (torch device tensors, torch's current stream) ---------
    def _torch_call(self, fn):
        """Runs fn(stream) ordered with torch's current stream.  A torch
        stream other than the default one is passed through, so the call
        stays asynchronous on it.  torch's default stream is the legacy null
        stream (handle 0), which this API reads as "the context's stream",
        and that stream is non-blocking, so the call is then fenced: torch's
        work is finished first and the context's after."""
        import torch
        s = torch.cuda.current_stream()
        if s.cuda_stream:
            return fn(s.cuda_stream)
        s.synchronize()
        out = fn(None)
        self.r.sync()
        return out

    def _tensor(self, t, shape, dtype, name):
        import torch
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()
                and tuple(t.shape) == tuple(shape)):
            raise ValueError(f"{name} must be a contiguous {dtype} device tensor of shape {tuple(shape)}")
        return t

    def aa_image(self, camera: N.Camera, out=None):
        """Antialiased linear image [H, W, 3] float32 of the last frame (the
        frame's flat resolve before the sRGB encode, blended along the front
        triangle's edge for every pixel pair of different colour)."""
        import torch
        if out is None:
            out = torch.empty((self.height, self.width, 3), dtype=torch.float32, device="cuda")
        self._tensor(out, (self.height, self.width, 3), torch.float32, "out")
        self._torch_call(lambda st: self.r._check(
            self.r._lib.vkgv_view_aa_image(self.r._h, self._h, C.byref(camera), out.data_ptr(), st)))
        return out

    def aa_image_noalloc(self, camera: N.Camera, out):
        """aa_image through vkgv_view_aa_image_noalloc: the library allocates
        nothing (its scratch is a fixed per-GPU array); `out` is required."""
        import torch
        self._tensor(out, (self.height, self.width, 3), torch.float32, "out")
        self._torch_call(lambda st: self.r._check(
            self.r._lib.vkgv_view_aa_image_noalloc(self.r._h, self._h, C.byref(camera), out.data_ptr(), st)))
        return out

    def aa_backward(self, camera: N.Camera, g_image, g_transforms):
        """g_transforms (float64 [T, 16] device tensor) += d(sum g_image * image) / d transforms."""
        import torch
        self._tensor(g_image, (self.height, self.width, 3), torch.float32, "g_image")
        if not (g_transforms.is_cuda and g_transforms.dtype == torch.float64 and g_transforms.is_contiguous()
                and g_transforms.dim() == 2 and g_transforms.shape[1] == 16):
            raise ValueError("g_transforms must be a contiguous float64 [T, 16] device tensor")
        self._torch_call(lambda st: self.r._check(self.r._lib.vkgv_view_aa_backward(
            self.r._h, self._h, C.byref(camera), g_image.data_ptr(), g_transforms.data_ptr(), g_transforms.shape[0],
            st)))

    def reset_history(self):
        self.r._check(self.r._lib.vkgv_view_reset_history(self.r._h, self._h))

    # -- readbacks ----------------------------------------------------------------
    def keys(self) -> np.ndarray:
        out = np.empty(self.width * self.height, dtype=np.uint64)
        self.r._check(self.r._lib.vkgv_view_read_visbuffer(self.r._h, self._h, out.ctypes.data))
        return out.reshape(self.height, self.width)

    def ids_depth(self):
        ids = np.empty(self.width * self.height, dtype=np.uint32)
        dep = np.empty(self.width * self.height, dtype=np.float32)
        self.r._check(self.r._lib.vkgv_view_read_ids_depth(self.r._h, self._h, ids.ctypes.data, dep.ctypes.data))
        return ids.reshape(self.height, self.width), dep.reshape(self.height, self.width)

    def motion(self) -> np.ndarray:
        """Motion vectors of the last VKGV_MOTION_VECTORS frame, (H, W, 2) float32."""
        out = np.empty(self.width * self.height * 2, dtype=np.float32)
        self.r._check(self.r._lib.vkgv_view_read_motion(self.r._h, self._h, out.ctypes.data))
        return out.reshape(self.height, self.width, 2)

    def color(self) -> np.ndarray:
        out = np.empty(self.width * self.height, dtype=np.uint32)
        self.r._check(self.r._lib.vkgv_view_read_color(self.r._h, self._h, out.ctypes.data))
        return out.reshape(self.height, self.width)

    def pyramid(self) -> np.ndarray:
        out = np.empty(self.pyramid_texels, dtype=np.float32)
        self.r._check(self.r._lib.vkgv_view_read_pyramid(self.r._h, self._h, out.ctypes.data))
        return out

    def pyramid_levels(self):
        flat = self.pyramid()
        res, off = [], 0
        for (w, h) in self.level_sizes:
            res.append(flat[off:off + w * h].reshape(h, w))
            off += w * h
        return res

    def write_pyramid(self, texels: np.ndarray):
        t = np.ascontiguousarray(texels, dtype=np.float32).ravel()
        if t.size != self.pyramid_texels:
            raise ValueError(f"pyramid has {self.pyramid_texels} texels, got {t.size}")
        self.r._check(self.r._lib.vkgv_view_write_pyramid(self.r._h, self._h, t.ctypes.data))

    def write_depth(self, depth: np.ndarray):
        d = np.ascontiguousarray(depth, dtype=np.float32).ravel()
        if d.size != self.width * self.height:
            raise ValueError("depth size mismatch")
        self.r._check(self.r._lib.vkgv_view_write_depth(self.r._h, self._h, d.ctypes.data))

    def visible(self, pass_index: int = 0) -> np.ndarray:
        n = C.c_uint32()
        self.r._check(self.r._lib.vkgv_view_read_visible(self.r._h, self._h, pass_index, None, 0, C.byref(n)))
        out = np.empty(max(1, n.value), dtype=np.uint32)
        self.r._check(self.r._lib.vkgv_view_read_visible(self.r._h, self._h, pass_index, out.ctypes.data,
                                                         out.size, C.byref(n)))
        return np.sort(out[:n.value])

    def stats(self) -> N.FrameStats:
        s = N.FrameStats()
        self.r._check(self.r._lib.vkgv_view_read_stats(self.r._h, self._h, C.byref(s)))
        return s
