"""glTF 2.0 scene ingestion (.gltf with external or data-URI buffers, .glb).

Follows the reference's asset loading (assets.cpp):
- positions through their accessor, with KHR_mesh_quantization normalization
  for normalized integer components (int8: max(c / 127, -1), uint8: c / 255,
  int16: max(c / 32767, -1), uint16: c / 65535) and integer values taken as
  they are when not normalized (fastgltf's iterateAccessor, assets.cpp:310-314);
- indices copied to u32, generated as 0..n-1 when a primitive has none
  (fastgltf::Options::GenerateMeshIndices, assets.cpp:533-534);
- the primitive AABB from the position accessor's min / max (assets.cpp:299-304);
- meshlets at the reference's limits (64 vertices, 124 triangles) by the
  host builder (assets.cpp:322-373);
- materials: index 0 is the default (albedo 1, alpha cutoff 0.5, single
  sided), glTF material i is i + 1 with its baseColorFactor, alphaCutoff and
  doubleSided (assets.cpp:478-518); a primitive's material index is its glTF
  index + 1, 0 without one (assets.cpp:292-294); a KHR_texture_transform of
  the base-colour texture is kept in the uv fields;
- beyond the reference (whose image load task is empty, assets.cpp:467-471,
  so its albedo handle is always invalid): base-colour textures decoded to
  RGBA8 with their sampler's wrap modes, the material's albedo_index pointing
  at them, and TEXCOORD uvs as float16 in the vertex, for the textured
  resolve;
- node matrices decomposed to translation / rotation / scale
  (fastgltf::Options::DecomposeNodeMatrices), then the World traversal
  (world.cpp:187-293, the host library's vkgv_world_build) of the default
  scene's roots gives the draws and transforms.
Beyond the reference, COLOR_0 (as unorm8 RGBA) is loaded into the vertex
colour that the barycentric resolve interpolates; without it the colour is
white.

EXT_meshopt_compression buffer views are decoded on the host
(vk_gltf_viewer_amd.meshopt, as assets.cpp:111-171 does with meshoptimizer)
before the accessors read them; a decoder error raises.

Not supported: primitive modes other than triangles (ValueError),
KHR_draco_mesh_compression (NotImplementedError).
"""
from __future__ import annotations

import base64
import ctypes as C
import json
import math
import os
import struct
import warnings

import numpy as np

from . import _native as N
from . import meshopt
from .scene import (ANIM_CHANNEL_DTYPE, ANIM_SAMPLER_DTYPE, DRAW_DTYPE, MATERIAL_DTYPE, NODE_DTYPE, PRIM_DTYPE,
                    Hierarchy, SceneArrays, Texture, build_meshlets)

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_COUNT = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_SUPPORTED_REQUIRED = {"EXT_meshopt_compression", "KHR_mesh_quantization", "KHR_texture_transform", "KHR_materials_emissive_strength",
                       "KHR_lights_punctual", "KHR_texture_basisu", "MSFT_texture_dds"}
INVALID_HANDLE = 0xFFFFFFFF


class GltfError(ValueError):
    pass


def _read_glb(data: bytes):
    magic, version, length = struct.unpack_from("<4sII", data, 0)
    if magic != b"glTF" or version != 2:
        raise GltfError("not a glTF 2.0 binary")
    off, js, binc = 12, None, None
    while off < min(length, len(data)):
        clen, ctype = struct.unpack_from("<I4s", data, off)
        chunk = data[off + 8:off + 8 + clen]
        if ctype == b"JSON":
            js = json.loads(chunk.decode("utf-8"))
        elif ctype == b"BIN\x00" and binc is None:
            binc = chunk
        off += 8 + clen
    if js is None:
        raise GltfError("glTF binary without a JSON chunk")
    return js, binc


class _Asset:
    def __init__(self, path: str):
        self.dir = os.path.dirname(os.path.abspath(path))
        with open(path, "rb") as f:
            data = f.read()
        if data[:4] == b"glTF":
            self.js, self.glb_bin = _read_glb(data)
        else:
            self.js, self.glb_bin = json.loads(data.decode("utf-8")), None
        if self.js.get("asset", {}).get("version", "2.0").split(".")[0] != "2":
            raise GltfError("only glTF 2.x is supported")
        required = set(self.js.get("extensionsRequired", []))
        unknown = required - _SUPPORTED_REQUIRED
        if unknown:
            raise GltfError(f"required extensions not supported: {sorted(unknown)}")
        self._buffers = {}
        self._decoded = {}

    def buffer(self, i: int) -> bytes:
        if i not in self._buffers:
            b = self.js["buffers"][i]
            uri = b.get("uri")
            if uri is None and b.get("extensions", {}).get("EXT_meshopt_compression", {}).get("fallback"):
                raise GltfError(f"buffer {i} is a meshopt fallback buffer without data")
            if uri is None:
                if i != 0 or self.glb_bin is None:
                    raise GltfError(f"buffer {i} has no uri and there is no GLB binary chunk")
                data = self.glb_bin
            elif uri.startswith("data:"):
                data = base64.b64decode(uri.split(",", 1)[1])
            else:
                from urllib.parse import unquote
                with open(os.path.join(self.dir, unquote(uri)), "rb") as f:
                    data = f.read()
            if len(data) < b["byteLength"]:
                raise GltfError(f"buffer {i} is shorter than its byteLength")
            self._buffers[i] = data
        return self._buffers[i]

    def view_bytes(self, vi: int):
        v = self.js["bufferViews"][vi]
        mc = v.get("extensions", {}).get("EXT_meshopt_compression")
        if mc is not None:
            if vi not in self._decoded:
                src = self.buffer(mc["buffer"])
                off = mc.get("byteOffset", 0)
                if off + mc["byteLength"] > len(src):
                    raise GltfError(f"buffer view {vi}: compressed range past the end of its buffer")
                self._decoded[vi] = meshopt.decode(src[off:off + mc["byteLength"]], mc["count"], mc["byteStride"],
                                                   mc["mode"], mc.get("filter", "NONE"))
            return self._decoded[vi], v.get("byteStride", mc["byteStride"] if mc["mode"] == "ATTRIBUTES" else None)
        data = self.buffer(v["buffer"])
        off = v.get("byteOffset", 0)
        return data[off:off + v["byteLength"]], v.get("byteStride")

    def accessor(self, ai: int, as_float: bool):
        """(count, components) array; with as_float, normalized integers are
        mapped as KHR_mesh_quantization / the core spec define."""
        a = self.js["accessors"][ai]
        dt = np.dtype(_COMPONENT[a["componentType"]]).newbyteorder("<")
        nc = _COUNT[a["type"]]
        count = a["count"]
        if "bufferView" in a:
            raw, stride = self.view_bytes(a["bufferView"])
            off = a.get("byteOffset", 0)
            elem = dt.itemsize * nc
            stride = stride or elem
            need = off + stride * (count - 1) + elem if count else 0
            if need > len(raw):
                raise GltfError(f"accessor {ai} reads past its buffer view")
            arr = np.ndarray((count, nc), dtype=dt, buffer=raw, offset=off, strides=(stride, dt.itemsize)).copy()
        else:
            arr = np.zeros((count, nc), dtype=dt)
        sp = a.get("sparse")
        if sp:
            idt = np.dtype(_COMPONENT[sp["indices"]["componentType"]]).newbyteorder("<")
            iraw, _ = self.view_bytes(sp["indices"]["bufferView"])
            idx = np.frombuffer(iraw, dtype=idt, count=sp["count"], offset=sp["indices"].get("byteOffset", 0))
            vraw, _ = self.view_bytes(sp["values"]["bufferView"])
            vals = np.frombuffer(vraw, dtype=dt, count=sp["count"] * nc,
                                 offset=sp["values"].get("byteOffset", 0)).reshape(-1, nc)
            if len(idx) and int(idx.max()) >= count:
                raise GltfError(f"accessor {ai}: sparse index out of range")
            arr[idx.astype(np.int64)] = vals
        if not as_float:
            return arr
        if a.get("normalized", False) and dt.kind in "iu":
            if dt == np.int8:
                return np.maximum(arr.astype(np.float32) / 127.0, -1.0).astype(np.float32)
            if dt == np.uint8:
                return (arr.astype(np.float32) / 255.0).astype(np.float32)
            if dt == np.int16:
                return np.maximum(arr.astype(np.float32) / 32767.0, -1.0).astype(np.float32)
            if dt == np.uint16:
                return (arr.astype(np.float32) / 65535.0).astype(np.float32)
            raise GltfError(f"accessor {ai}: normalized {dt} is not allowed")
        return arr.astype(np.float32)


def _quat_from_matrix(R):
    """Unit quaternion (x, y, z, w) of a rotation matrix (Shepperd's method)."""
    t = R[0, 0] + R[1, 1] + R[2, 2]
    if t > 0:
        s = math.sqrt(t + 1.0) * 2
        w, x, y, z = 0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        w, x, y, z = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        w, x, y, z = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        w, x, y, z = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    return q / np.linalg.norm(q)


def decompose(matrix16):
    """glTF column-major node matrix -> (translation, rotation xyzw, scale),
    for matrices without shear (as glTF requires of node matrices); a negative
    determinant is carried by the x scale."""
    M = np.asarray(matrix16, np.float64).reshape(4, 4).T
    t = M[:3, 3].copy()
    A = M[:3, :3].copy()
    s = np.linalg.norm(A, axis=0)
    if np.linalg.det(A) < 0:
        s[0] = -s[0]
    R = A / np.where(s == 0, 1.0, s)
    return t, _quat_from_matrix(R), s


_WRAP = {10497: 0, 33071: 1, 33648: 2}     # REPEAT, CLAMP_TO_EDGE, MIRRORED_REPEAT


def _image_bytes(asset, img):
    if "bufferView" in img:
        raw, _ = asset.view_bytes(img["bufferView"])
        return bytes(raw)
    uri = img.get("uri", "")
    if uri.startswith("data:"):
        return base64.b64decode(uri.split(",", 1)[1])
    from urllib.parse import unquote
    with open(os.path.join(asset.dir, unquote(uri)), "rb") as f:
        return f.read()


def _load_textures(asset):
    """One Texture per glTF texture (image + sampler), None where the image
    cannot be decoded (no PIL, or a format PIL does not read, such as KTX2)."""
    js = asset.js
    out, cache = [], {}
    try:
        from PIL import Image
    except ImportError:                      # pragma: no cover - PIL is part of the image
        Image = None
    for ti, t in enumerate(js.get("textures", [])):
        src = t.get("source")
        for ext in ("KHR_texture_basisu", "MSFT_texture_dds"):
            if src is None and ext in t.get("extensions", {}):
                src = t["extensions"][ext].get("source")
        tex = None
        if src is not None and Image is not None:
            if src not in cache:
                try:
                    import io
                    with Image.open(io.BytesIO(_image_bytes(asset, js["images"][src]))) as im:
                        cache[src] = np.ascontiguousarray(np.asarray(im.convert("RGBA"), np.uint8))
                except Exception as e:          # undecodable image: the material draws untextured
                    warnings.warn(f"glTF image {src} not decoded ({e}); texture {ti} is skipped")
                    cache[src] = None
            rgba = cache[src]
            if rgba is not None:
                smp = js["samplers"][t["sampler"]] if "sampler" in t else {}
                tex = Texture(rgba, _WRAP.get(smp.get("wrapS", 10497), 0), _WRAP.get(smp.get("wrapT", 10497), 0))
        out.append(tex)
    return out


def load_gltf(path: str, scene: int | None = None, textures: bool = True) -> SceneArrays:
    """The glTF file at `path` as SceneArrays (meshes, primitives, materials,
    draws and transforms of the default scene, or of `scene`).  With
    `textures`, base-colour textures are decoded (PNG, JPEG and whatever else
    PIL reads) and the materials' albedo_index points at them; each
    primitive's vertex uv is the TEXCOORD set its base-colour texture uses."""
    asset = _Asset(path)
    js = asset.js
    gtex = _load_textures(asset) if textures else []
    tex_list, tex_map = [], {}
    # materials: 0 = default, glTF i -> i + 1
    gmats = js.get("materials", [])
    mats = np.zeros(len(gmats) + 1, MATERIAL_DTYPE)
    mats["albedo_factor"] = 1.0
    mats["albedo_index"] = INVALID_HANDLE
    mats["uv_scale"] = 1.0
    mats["alpha_cutoff"] = 0.5
    for i, m in enumerate(gmats):
        pbr = m.get("pbrMetallicRoughness", {})
        mats[i + 1]["albedo_factor"] = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        mats[i + 1]["alpha_cutoff"] = m.get("alphaCutoff", 0.5)
        mats[i + 1]["double_sided"] = 1 if m.get("doubleSided", False) else 0
        bct = pbr.get("baseColorTexture")
        if bct is not None and bct.get("index", -1) < len(gtex) and gtex[bct["index"]] is not None:
            gi = bct["index"]
            if gi not in tex_map:
                tex_map[gi] = len(tex_list)
                tex_list.append(gtex[gi])
            mats[i + 1]["albedo_index"] = tex_map[gi]
        tt = (bct or {}).get("extensions", {}).get("KHR_texture_transform")
        if tt:
            mats[i + 1]["uv_offset"] = tt.get("offset", [0.0, 0.0])
            mats[i + 1]["uv_scale"] = tt.get("scale", [1.0, 1.0])
            mats[i + 1]["uv_rotation"] = tt.get("rotation", 0.0)
    # meshes -> one vkgv mesh + primitive per glTF primitive
    meshes, prims, mesh_prims, prim_list = [], [], [], []
    for gm in js.get("meshes", []):
        first = len(prims)
        for gp in gm["primitives"]:
            if gp.get("mode", 4) != 4:
                raise ValueError("only triangle-list primitives are supported (mode 4)")
            if "extensions" in gp and "KHR_draco_mesh_compression" in gp["extensions"]:
                raise NotImplementedError("KHR_draco_mesh_compression is not supported")
            pa = gp["attributes"]["POSITION"]
            pos = asset.accessor(pa, True)
            if pos.shape[1] != 3:
                raise GltfError("POSITION must be VEC3")
            if "indices" in gp:
                idx = asset.accessor(gp["indices"], False).ravel().astype(np.uint32)
            else:
                idx = np.arange(pos.shape[0], dtype=np.uint32)
            if len(idx) % 3:
                raise GltfError("triangle list with an index count not divisible by 3")
            if len(idx) and int(idx.max()) >= pos.shape[0]:
                raise GltfError("index out of range")
            ma = build_meshlets(pos, idx)
            uv_set = _uv_set(js, gp)
            if f"TEXCOORD_{uv_set}" in gp["attributes"]:
                uv = asset.accessor(gp["attributes"][f"TEXCOORD_{uv_set}"], True)
                if uv.shape != (pos.shape[0], 2):
                    raise GltfError("TEXCOORD must be VEC2 with one entry per vertex")
                ma.vertices["uv"] = uv.astype(np.float16).view(np.uint16)
            if "COLOR_0" in gp["attributes"]:
                col = asset.accessor(gp["attributes"]["COLOR_0"], True)
                rgba = np.ones((col.shape[0], 4), np.float32)
                rgba[:, :col.shape[1]] = col
                ma.vertices["color"] = np.clip(np.rint(rgba * 255.0), 0, 255).astype(np.uint8)
            acc = js["accessors"][pa]
            if "min" in acc and "max" in acc:
                lo, hi = np.asarray(acc["min"], np.float32), np.asarray(acc["max"], np.float32)
                if acc.get("normalized", False):
                    lo = asset_normalize(acc, lo)
                    hi = asset_normalize(acc, hi)
            else:
                lo, hi = pos.min(0), pos.max(0)
            p = np.zeros(1, PRIM_DTYPE)[0]
            p["mesh_index"] = len(meshes)
            p["material_index"] = gp["material"] + 1 if "material" in gp else 0
            center = (lo + hi) / 2.0
            p["aabb_center"] = center
            p["aabb_extents"] = hi - center
            meshes.append(ma)
            prims.append(p)
            prim_list.append(len(prims) - 1)
        mesh_prims += [first, len(prims) - first]
    prims_arr = np.array(prims, dtype=PRIM_DTYPE) if prims else np.zeros(0, PRIM_DTYPE)
    # nodes of the scene -> draws and transforms (World traversal)
    gnodes = js.get("nodes", [])
    nodes = np.zeros(len(gnodes), NODE_DTYPE)
    children = []
    for i, n in enumerate(gnodes):
        kids = n.get("children", [])
        nodes[i]["mesh"] = n.get("mesh", -1)
        nodes[i]["first_child"], nodes[i]["child_count"] = len(children), len(kids)
        children += kids
        if "matrix" in n:
            t, r, s = decompose(n["matrix"])
        else:
            t = n.get("translation", [0.0, 0.0, 0.0])
            r = n.get("rotation", [0.0, 0.0, 0.0, 1.0])
            s = n.get("scale", [1.0, 1.0, 1.0])
        nodes[i]["translation"], nodes[i]["rotation"], nodes[i]["scale"] = t, r, s
    scenes = js.get("scenes", [])
    si = scene if scene is not None else js.get("scene", 0)
    # without a scene the reference draws nothing (world.cpp:233-234)
    roots = scenes[si].get("nodes", []) if scenes else []
    draws, transforms = _world(nodes, children, roots, mesh_prims, prim_list, [len(m.meshlets) for m in meshes])
    hier = Hierarchy(nodes, np.asarray(children, np.uint32), np.asarray(roots, np.uint32),
                     *_animations(asset, len(gnodes)))
    return SceneArrays(meshes, prims_arr, mats, draws, transforms, textures=tex_list, hierarchy=hier)


def _uv_set(js, gp):
    """TEXCOORD set of the primitive's base-colour texture (0 by default;
    KHR_texture_transform may override it)."""
    if "material" not in gp:
        return 0
    bct = js["materials"][gp["material"]].get("pbrMetallicRoughness", {}).get("baseColorTexture", {})
    tt = bct.get("extensions", {}).get("KHR_texture_transform", {})
    return int(tt.get("texCoord", bct.get("texCoord", 0)))


def asset_normalize(acc, v):
    """Accessor min / max of a normalized integer accessor, in float."""
    ct = acc["componentType"]
    if ct == 5120:
        return np.maximum(v / 127.0, -1.0).astype(np.float32)
    if ct == 5121:
        return (v / 255.0).astype(np.float32)
    if ct == 5122:
        return np.maximum(v / 32767.0, -1.0).astype(np.float32)
    if ct == 5123:
        return (v / 65535.0).astype(np.float32)
    return v


_INTERP = {"LINEAR": 0, "STEP": 1, "CUBICSPLINE": 2}
_PATH = {"translation": 0, "rotation": 1, "scale": 2}


def _animations(asset, node_count):
    """All animations' samplers and channels, flattened in animation order
    (assets.cpp:579-600): keyframe times and output values as floats
    (normalized integer outputs decoded), channels of unsupported paths
    (morph-target weights) dropped as the reference does (world.cpp:214-215)."""
    js = asset.js
    samplers, channels, times, values = [], [], [], []
    nt = nv = 0
    for anim in js.get("animations", []):
        base = len(samplers)
        for sp in anim.get("samplers", []):
            t = asset.accessor(sp["input"], True).ravel()
            v = asset.accessor(sp["output"], True).ravel()
            if len(t) == 0:
                raise GltfError("animation sampler without keyframes")
            interp = _INTERP.get(sp.get("interpolation", "LINEAR"))
            if interp is None:
                raise GltfError(f"unknown interpolation {sp.get('interpolation')}")
            samplers.append((nt, len(t), nv, interp))
            times.append(t)
            values.append(v)
            nt += len(t)
            nv += len(v)
        for ch in anim.get("channels", []):
            path = _PATH.get(ch["target"].get("path"))
            node = ch["target"].get("node")
            if path is None or node is None:
                continue
            if node >= node_count:
                raise GltfError("animation channel targets a missing node")
            channels.append((node, base + ch["sampler"], path, 0))
    sp = np.array(samplers, dtype=np.uint32).reshape(-1, 4).view(ANIM_SAMPLER_DTYPE).ravel()
    cn = np.array(channels, dtype=np.uint32).reshape(-1, 4).view(ANIM_CHANNEL_DTYPE).ravel()
    tm = np.concatenate(times).astype(np.float32) if times else np.zeros(0, np.float32)
    vl = np.concatenate(values).astype(np.float32) if values else np.zeros(0, np.float32)
    return sp, tm, vl, cn


def _world(nodes, children, roots, mesh_prims, prim_list, meshlet_counts):
    L = N.host()
    nd_arr = np.ascontiguousarray(nodes if len(nodes) else np.zeros(1, NODE_DTYPE))
    ch = np.ascontiguousarray(children if children else [0], np.uint32)
    rt = np.ascontiguousarray(roots if roots else [0], np.uint32)
    mp = np.ascontiguousarray(mesh_prims if mesh_prims else [0, 0], np.uint32)
    pl = np.ascontiguousarray(prim_list if prim_list else [0], np.uint32)
    mc = np.ascontiguousarray(meshlet_counts if meshlet_counts else [0], np.uint32)
    args = [nd_arr.ctypes.data_as(C.POINTER(N.Node)), len(nodes), ch.ctypes.data, rt.ctypes.data, len(roots),
            mp.ctypes.data, pl.ctypes.data, mc.ctypes.data]
    nd, nt = C.c_uint32(0), C.c_uint32(0)
    if L.vkgv_world_build(*args, None, C.byref(nd), None, C.byref(nt)) != 0:
        raise GltfError("World traversal failed: " + L.vkgv_scene_error().decode())
    draws = np.zeros(nd.value, DRAW_DTYPE)
    tr = np.zeros((nt.value, 16), np.float32)
    if L.vkgv_world_build(*args, draws.ctypes.data, C.byref(nd), tr.ctypes.data, C.byref(nt)) != 0:
        raise GltfError("World traversal failed: " + L.vkgv_scene_error().decode())
    return draws, tr
