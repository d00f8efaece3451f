"""Small scenes built from explicit geometry, for tests."""
from __future__ import annotations

import numpy as np

from vk_gltf_viewer_amd.scene import (DRAW_DTYPE, MATERIAL_DTYPE, PRIM_DTYPE, SceneArrays, build_meshlets)


def colmajor(m):
    """Row-major 4x4 -> column-major float[16]."""
    return np.asarray(m, np.float32).T.ravel()


def tiny_scene(meshes, instances, colors=None) -> SceneArrays:
    """meshes: [(positions (V,3), indices (T*3,))]; instances: [(mesh index,
    row-major 4x4 transform, rgba albedo, double_sided)]; colors: optional
    per-mesh (V,4) uint8 vertex colours.  One primitive and one material per
    instance; every meshlet of the instance is a draw."""
    built = []
    for i, (pos, idx) in enumerate(meshes):
        m = build_meshlets(np.asarray(pos, np.float32), np.asarray(idx, np.uint32))
        if colors is not None and colors[i] is not None:
            m.vertices["color"] = colors[i]
        built.append(m)
    prims = np.zeros(len(instances), PRIM_DTYPE)
    mats = np.zeros(len(instances), MATERIAL_DTYPE)
    tr = np.zeros((len(instances), 16), np.float32)
    draws = []
    for k, (mi, M, albedo, ds) in enumerate(instances):
        prims[k]["mesh_index"] = mi
        prims[k]["material_index"] = k
        mats[k]["albedo_factor"] = albedo
        mats[k]["uv_scale"] = (1, 1)
        mats[k]["double_sided"] = 1 if ds else 0
        tr[k] = colmajor(M)
        for ml in range(len(built[mi].meshlets)):
            draws.append((k, ml, k))
    d = np.array(draws, dtype=np.uint32).reshape(-1, 3)
    dr = np.zeros(len(d), DRAW_DTYPE)
    dr["primitive_index"], dr["meshlet_index"], dr["transform_index"] = d[:, 0], d[:, 1], d[:, 2]
    return SceneArrays(built, prims, mats, dr, tr)


def icosphere(level=2):
    """Unit icosphere, counter-clockwise seen from outside."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        for (a, b, c) in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts, np.float32), np.array(f, np.uint32).ravel()


def translate_scale(t, s=1.0):
    M = np.eye(4)
    M[:3, :3] *= s
    M[:3, 3] = t
    return M
