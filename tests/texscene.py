"""Scenes for the textured resolve: a screen-filling quad whose vertex uvs
put pixel centres on texel centres, and a textured glTF asset."""
import io

import numpy as np

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import (DRAW_DTYPE, MATERIAL_DTYPE, PRIM_DTYPE, SceneArrays, Texture, build_meshlets,
                                      camera_from_view, view_desc)


def quad_scene(W, H, texture: Texture, uv_offset=(0.0, 0.0), uv_scale=(1.0, 1.0), uv_rotation=0.0,
               albedo=(1.0, 1.0, 1.0, 1.0)):
    """An orthographic W x H view of a quad spanning NDC [-1.5, 1.5]^2; the
    vertex uvs are (ndc + 1) / 2 (exact in float16), so the interpolated uv
    at pixel (px, py) is ((px + 0.5) / W, (py + 0.5) / H) and, with an
    identity uv transform, lands on texel (px, py) of a W x H texture."""
    vd = view_desc(W, H, (0.0, 0.0, 2.0), (0.0, 0.0, 0.0), znear=0.5, zfar=4.0, ortho_half=(1.0, 1.0))
    cam = camera_from_view(vd)
    pos = np.array([[-1.5, -1.5, 0.0], [1.5, -1.5, 0.0], [-1.5, 1.5, 0.0], [1.5, 1.5, 0.0]], np.float32)
    idx = np.array([0, 1, 2, 2, 1, 3], np.uint32)
    m = build_meshlets(pos, idx)
    vp = np.asarray(cam.view_projection, np.float64).reshape(4, 4).T
    clip = (vp @ np.c_[pos, np.ones(4)].T).T
    ndc = clip[:, :2] / clip[:, 3:4]
    uv = (ndc + 1.0) / 2.0
    assert np.allclose(np.abs(ndc), 1.5, atol=1e-6)
    m.vertices["uv"] = (np.rint(uv * 4) / 4).astype(np.float16).view(np.uint16)   # -0.25 or 1.25
    prim = np.zeros(1, PRIM_DTYPE)
    prim["aabb_extents"] = (1.5, 1.5, 0.0)
    mat = np.zeros(1, MATERIAL_DTYPE)
    mat["albedo_factor"] = albedo
    mat["albedo_index"] = 0
    mat["uv_offset"] = uv_offset
    mat["uv_scale"] = uv_scale
    mat["uv_rotation"] = uv_rotation
    mat["alpha_cutoff"] = 0.5
    mat["double_sided"] = 1
    draws = np.zeros(len(m.meshlets), DRAW_DTYPE)
    draws["meshlet_index"] = np.arange(len(m.meshlets))
    sc = SceneArrays([m], prim, mat, draws, np.eye(4, dtype=np.float32).reshape(1, 16), views=[vd],
                     textures=[texture])
    return sc, cam


def png_bytes(rgba):
    from PIL import Image
    buf = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(rgba, np.uint8), "RGBA").save(buf, format="PNG")
    return buf.getvalue()


def textured_spheres(grid=6, seed=3):
    """glTF builder: a grid of uv spheres with TEXCOORD_0 and PNG base-colour
    textures under REPEAT, MIRRORED_REPEAT and CLAMP_TO_EDGE samplers, one
    with KHR_texture_transform, plus an untextured material."""
    from gltfgen import Builder, rotation_quat, uv_sphere
    rng = np.random.default_rng(seed)
    b = Builder()
    pos, tris = uv_sphere(24, 32)
    # uvs: longitude and latitude, spanning [0, 2] x [0, 1] so that the wrap modes matter
    lon = (np.arctan2(pos[:, 2], pos[:, 0]) / np.pi + 1.0)
    lat = np.arccos(np.clip(pos[:, 1], -1, 1)) / np.pi
    uv = np.c_[lon, lat].astype(np.float32)
    pa = b.accessor(pos, minmax=True)
    ta = b.accessor(uv)
    ia = b.accessor(tris.astype(np.uint16).reshape(-1, 1))
    b.js["images"], b.js["textures"], b.js["samplers"] = [], [], []
    for k, wrap in enumerate((10497, 33648, 33071)):
        img = (rng.integers(0, 256, (16 + 8 * k, 24, 4))).astype(np.uint8)
        img[..., 3] = 255 - 40 * k
        b.js["images"].append({"bufferView": b.view(png_bytes(img)), "mimeType": "image/png"})
        b.js["samplers"].append({"wrapS": wrap, "wrapT": 33071 if wrap != 33071 else 10497})
        b.js["textures"].append({"source": k, "sampler": k})
    mats = []
    for k in range(3):
        extra = {"baseColorTexture": {"index": k}}
        if k == 1:
            extra["baseColorTexture"]["extensions"] = {"KHR_texture_transform": {
                "offset": [0.25, -0.5], "scale": [1.5, 0.75], "rotation": 0.3}}
        mats.append(b.material(rng.uniform(0.4, 1.0, 3).tolist() + [1.0], extra=extra))
    mats.append(b.material([0.8, 0.3, 0.2, 1.0]))
    meshes = [b.mesh([{"attributes": {"POSITION": pa, "TEXCOORD_0": ta}, "indices": ia, "material": m}])
              for m in mats]
    kids = []
    for i in range(grid):
        for j in range(grid):
            kids.append(b.node(mesh=int(rng.integers(0, len(meshes))),
                               translation=[float((i - grid / 2) * 2.3), 0.0, float((j - grid / 2) * 2.3)],
                               rotation=rotation_quat([0, 1, 0], rng.uniform(0, 6.28)).tolist()))
    b.scene(kids)
    return b
