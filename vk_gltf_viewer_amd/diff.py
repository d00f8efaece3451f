"""Differentiable form of the frame (torch).

    image = antialias(interpolate(clip, attr, ids), clip, ids)
    clip  = VP_c @ M_i @ [p_v, 1]        (torch: gradients to positions and transforms)
    ids   = rasterize(clip)              (no gradient)

`rasterize`, `interpolate` and `antialias` are the LibTorch extension
`vkgv_diff` (lib/vkgv_diff.so over lib/libvkgv_diffk.so, HIP kernels for
gfx950): interpolation and antialiasing are torch.autograd Functions whose
backward kernels recompute barycentrics and edge crossings from the id
buffer.  The definitions are in vk_gltf_viewer_amd/csrc/diff/diff.hip.
"""
from __future__ import annotations

import ctypes
import importlib.machinery
import importlib.util
import math
import os
from dataclasses import dataclass

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_ext = None


def ext():
    """The extension module; raises if it is not built (no fallback)."""
    global _ext
    if _ext is None:
        import torch  # noqa: F401  (the extension links against the loaded libtorch)
        kern = os.path.join(_LIB, "libvkgv_diffk.so")
        mod = os.path.join(_LIB, "vkgv_diff.so")
        if not (os.path.exists(kern) and os.path.exists(mod)):
            raise RuntimeError(f"{mod} is not built: run __graft_entry__.build()")
        ctypes.CDLL(kern, mode=ctypes.RTLD_GLOBAL)
        loader = importlib.machinery.ExtensionFileLoader("vkgv_diff", mod)
        spec = importlib.util.spec_from_file_location("vkgv_diff", mod, loader=loader)
        m = importlib.util.module_from_spec(spec)
        loader.exec_module(m)
        _ext = m
    return _ext


def clip_positions(pos, xforms, view_projs):
    """pos [V, 3], xforms [I, 4, 4], view_projs [C, 4, 4] (row-major, column
    vectors) -> clip [C, I, V, 4], by the extension's vertex-transform
    kernels (gradients to positions and to both matrix sets)."""
    A = (view_projs[:, None, :, :, None] * xforms[None, :, None, :, :]).sum(-2)   # VP_c M_i, [C, I, 4, 4]
    return ext().transform(pos.contiguous(), A.contiguous())


PRECISIONS = ("float32", "mixed", "float64")


def render(pos, attr, tris, xforms, view_projs, height: int, width: int, antialias: bool = True,
           precision: str = "mixed"):
    """(image [C, H, W, A], ids [C, H, W], depth [C, H, W]).

    precision: the types of the clip positions and of the interpolation and
    antialiasing (forward and backward):
      float32  clip positions and arithmetic in float32;
      mixed    clip positions in float64, the pixel-relative offsets formed
               in float64, images, attributes and the rest in float32;
      float64  everything in float64.
    The antialiasing edge term is ill-conditioned in the clip positions (it
    scales as 1/den^2 for an edge nearly parallel to a pixel pair), so their
    float32 rounding alone moves the exact gradient by about 1e-4; "mixed"
    removes that at float32 cost.  Gradients arrive in the parameters' own
    type.  The raster always takes the clip positions rounded to float32."""
    e = ext()
    import torch
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {PRECISIONS}")
    cdt = torch.float32 if precision == "float32" else torch.float64
    adt = torch.float64 if precision == "float64" else torch.float32
    clip = clip_positions(pos.to(cdt), xforms.to(cdt), view_projs.to(cdt))
    ids, depth = e.rasterize(clip.detach().float().contiguous(), tris, height, width, False)
    attr = attr.to(adt).contiguous()
    if antialias:   # interpolation and antialiasing as one function: one backward vertex pass
        img = e.shade(clip, attr, tris, ids, depth)
    else:
        img = e.interpolate(clip, attr, tris, ids)
    return img, ids, depth


@dataclass
class DiffScene:
    pos: np.ndarray          # [V, 3] float32
    tris: np.ndarray         # [T, 3] int32
    albedo: np.ndarray       # [V, 3] float32, linear
    xforms: np.ndarray       # [I, 4, 4] float32
    view_projs: np.ndarray   # [C, 4, 4] float32
    height: int
    width: int


def camera_matrix(view) -> np.ndarray:
    """Row-major view-projection of a view (the frame's reverse-Z camera)."""
    from .scene import camera_from_view
    cam = camera_from_view(view)
    return np.asarray(cam.view_projection, np.float32).reshape(4, 4).T.copy()


def c4_scene(seed: int = 4, instances: int = 8, cameras: int = 8, size: int = 512) -> DiffScene:
    """The differentiable configuration: the subdivision-6 icosphere (40,962
    vertices, 81,920 triangles, radius 1) with its per-vertex albedo, as
    `instances` copies on a circle of radius 1.1 (scale 0.45, seeded rotation
    about y; one instance has the identity transform), seen by `cameras`
    cameras on a ring of radius 3 at size x size."""
    from .scene import Scene, view_desc
    sc = Scene(4, seed).arrays()
    m = sc.meshes[0]
    tris = []
    for ml in m.meshlets:
        vo, to, tc = int(ml["vertex_offset"]), int(ml["triangle_offset"]), int(ml["triangle_count"])
        loc = m.triangle_indices[to:to + 3 * tc].reshape(-1, 3).astype(np.int64)
        tris.append(m.vertex_indices[vo + loc])
    tris = np.concatenate(tris).astype(np.int32)
    pos = m.vertices["position"].astype(np.float32)
    albedo = (m.vertices["color"][:, :3].astype(np.float32) / 255.0) ** 2.2
    rng = np.random.default_rng(seed)
    xf = []
    for k in range(instances):
        M = np.eye(4, dtype=np.float32)
        if instances > 1:
            a = 2 * math.pi * k / instances
            r = rng.uniform(0, 2 * math.pi)
            M[:3, :3] = 0.45 * np.array([[math.cos(r), 0, math.sin(r)], [0, 1, 0], [-math.sin(r), 0, math.cos(r)]])
            M[:3, 3] = [1.1 * math.cos(a), 0.15 * math.sin(3 * a), 1.1 * math.sin(a)]
        xf.append(M)
    vps = []
    for k in range(cameras):
        a = 2 * math.pi * (k + 0.5) / cameras
        v = view_desc(size, size, (3.0 * math.cos(a), 0.8, 3.0 * math.sin(a)), (0.0, 0.0, 0.0), fovy_deg=45.0)
        vps.append(camera_matrix(v))
    return DiffScene(pos, tris, albedo, np.stack(xf), np.stack(vps), size, size)


def c4_target(scene: DiffScene, seed: int = 44):
    """The target of the configuration's L2 loss: the same mesh with
    N(0, 0.01) added to every vertex coordinate and N(0, 0.05) to every
    albedo channel (positions [V, 3], albedo [V, 3], float32)."""
    rng = np.random.default_rng(seed)
    pos = scene.pos + rng.normal(0.0, 0.01, scene.pos.shape).astype(np.float32)
    alb = scene.albedo + rng.normal(0.0, 0.05, scene.albedo.shape).astype(np.float32)
    return pos, alb
