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
import functools
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


@functools.lru_cache(maxsize=4)
def _geodesic_sphere(n: int):
    return _build_geodesic_sphere(n)


def geodesic_sphere(frequency: int):
    """Unit geodesic icosphere of class I (copies of a cached build): see
    _build_geodesic_sphere."""
    pos, tris = _geodesic_sphere(int(frequency))
    return pos.copy(), tris.copy()


def _build_geodesic_sphere(frequency: int):
    """Unit geodesic icosphere of class I: each icosahedron face cut into
    frequency^2 triangles on its barycentric grid, the grid points projected
    to the sphere.  10 f^2 + 2 vertices, 20 f^2 triangles, outward
    counter-clockwise winding (a face's grid triangles keep its orientation).
    Returns (pos [V, 3] float64, tris [T, 3] int32)."""
    n = int(frequency)
    t = (1.0 + 5.0 ** 0.5) / 2.0
    ico = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                    [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    index = {}
    pos = []

    def vertex(weights):
        # a grid point is sum w_k ico[k] / n with integer weights; shared
        # edge and corner points get one key from their nonzero weights
        key = tuple(sorted((k, w) for k, w in weights.items() if w))
        v = index.get(key)
        if v is None:
            p = sum(w * ico[k] for k, w in key) / n
            v = index[key] = len(pos)
            pos.append(p / np.linalg.norm(p))
        return v

    tris = []
    for a, b, c in faces:
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                w = {}
                for k, x in ((a, n - i - j), (b, i), (c, j)):
                    w[k] = w.get(k, 0) + x
                grid[i, j] = vertex(w)
        for i in range(n):
            for j in range(n - i):
                tris.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j < n - 1:
                    tris.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    return np.array(pos), np.array(tris, np.int32)


def c4_scene(seed: int = 4, instances: int = 8, cameras: int = 8, size: int = 512,
             frequency: int = 71) -> DiffScene:
    """The differentiable configuration: a 50k-vertex icosphere-derived mesh,
    the frequency-71 geodesic icosphere (50,412 vertices, 100,820 triangles,
    radius 1), with a seeded smooth per-vertex linear RGB albedo in
    [0.1, 0.9], as `instances` copies on a circle of radius 1.1 (scale 0.45,
    seeded rotation about y; one instance has the identity transform), seen
    by `cameras` cameras on a ring of radius 3 at size x size."""
    from .scene import view_desc
    p64, tris = geodesic_sphere(frequency)
    pos = p64.astype(np.float32)
    ph = np.random.default_rng(seed + 1000).uniform(0.0, 2.0 * math.pi, (3, 3))
    albedo = np.stack([0.5 + 0.4 * np.sin(3.0 * p64[:, 0] + ph[c, 0]) * np.cos(2.0 * p64[:, 1] + ph[c, 1])
                       * np.cos(2.5 * p64[:, 2] + ph[c, 2]) for c in range(3)], 1).astype(np.float32)
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
