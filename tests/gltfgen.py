"""Writes small glTF 2.0 assets for the loader tests (.gltf with a data-URI or
external buffer, .glb), from numpy arrays, without any glTF library."""
import base64
import json
import struct

import numpy as np

CT = {np.dtype(np.int8): 5120, np.dtype(np.uint8): 5121, np.dtype(np.int16): 5122,
      np.dtype(np.uint16): 5123, np.dtype(np.uint32): 5125, np.dtype(np.float32): 5126}
TYPE = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}


class Builder:
    def __init__(self):
        self.bin = bytearray()
        self.fallback = 0          # byteLength of the meshopt fallback buffer (buffer 1), 0 = none
        self.js = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [],
                   "meshes": [], "materials": [], "nodes": [], "scenes": [], "scene": 0}

    def _align(self, n=4):
        while len(self.bin) % n:
            self.bin.append(0)

    def view(self, data: bytes, stride=None, target=None):
        self._align()
        v = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        if target:
            v["target"] = target
        self.bin += data
        self.js["bufferViews"].append(v)
        return len(self.js["bufferViews"]) - 1

    def compressed_view(self, stream: bytes, count, stride, mode, filter=None, target=None):
        """An EXT_meshopt_compression buffer view: `stream` goes into buffer 0,
        the view itself points into the fallback buffer 1."""
        self._align()
        ext = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(stream), "byteStride": stride,
               "count": count, "mode": mode}
        if filter:
            ext["filter"] = filter
        self.bin += stream
        v = {"buffer": 1, "byteOffset": self.fallback, "byteLength": count * stride,
             "extensions": {"EXT_meshopt_compression": ext}}
        if mode == "ATTRIBUTES":
            v["byteStride"] = stride
        if target:
            v["target"] = target
        self.fallback += (count * stride + 3) & ~3
        self.js["bufferViews"].append(v)
        for k in ("extensionsUsed", "extensionsRequired"):
            if "EXT_meshopt_compression" not in self.js.setdefault(k, []):
                self.js[k].append("EXT_meshopt_compression")
        return len(self.js["bufferViews"]) - 1

    def accessor(self, arr, normalized=False, minmax=False, view=None, offset=0, count=None):
        arr = np.asarray(arr)
        a2 = arr.reshape(len(arr), -1)
        if view is None:
            view = self.view(np.ascontiguousarray(a2).tobytes())
        acc = {"bufferView": view, "componentType": CT[a2.dtype], "count": count if count is not None else len(a2),
               "type": TYPE[a2.shape[1]]}
        if offset:
            acc["byteOffset"] = offset
        if normalized:
            acc["normalized"] = True
        if minmax:
            acc["min"] = [a2[:, i].min().item() for i in range(a2.shape[1])]
            acc["max"] = [a2[:, i].max().item() for i in range(a2.shape[1])]
        self.js["accessors"].append(acc)
        return len(self.js["accessors"]) - 1

    def mesh(self, prims):
        self.js["meshes"].append({"primitives": prims})
        return len(self.js["meshes"]) - 1

    def material(self, color, cutoff=None, double_sided=False, extra=None):
        m = {"pbrMetallicRoughness": {"baseColorFactor": list(map(float, color))}}
        if cutoff is not None:
            m["alphaCutoff"] = cutoff
        if double_sided:
            m["doubleSided"] = True
        if extra:
            m["pbrMetallicRoughness"].update(extra)
        self.js["materials"].append(m)
        return len(self.js["materials"]) - 1

    def node(self, **kw):
        self.js["nodes"].append(kw)
        return len(self.js["nodes"]) - 1

    def scene(self, roots):
        self.js["scenes"].append({"nodes": list(roots)})

    def _json(self, uri):
        js = json.loads(json.dumps(self.js))
        for k in ("materials", "nodes", "meshes", "scenes"):
            if not js[k]:
                del js[k]
        if "scenes" not in js:
            js.pop("scene", None)
        self._align()
        b = {"byteLength": len(self.bin)}
        if uri is not None:
            b["uri"] = uri
        js["buffers"] = [b]
        if self.fallback:
            js["buffers"].append({"byteLength": self.fallback,
                                  "extensions": {"EXT_meshopt_compression": {"fallback": True}}})
        return js

    def write_gltf(self, path, external=False):
        self._align()
        if external:
            binpath = str(path) + ".bin"
            with open(binpath, "wb") as f:
                f.write(bytes(self.bin))
            uri = binpath.rsplit("/", 1)[-1]
        else:
            uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        with open(path, "w") as f:
            json.dump(self._json(uri), f)

    def write_glb(self, path):
        self._align()
        j = json.dumps(self._json(None)).encode()
        j += b" " * ((4 - len(j) % 4) % 4)
        b = bytes(self.bin)
        total = 12 + 8 + len(j) + 8 + len(b)
        with open(path, "wb") as f:
            f.write(struct.pack("<4sII", b"glTF", 2, total))
            f.write(struct.pack("<I4s", len(j), b"JSON") + j)
            f.write(struct.pack("<I4s", len(b), b"BIN\x00") + b)


def uv_sphere(rings, segments):
    """Closed sphere of radius 1: (V, 3) float32 positions, (T, 3) u32 triangles."""
    pos = [(0.0, 1.0, 0.0)]
    for r in range(1, rings):
        th = np.pi * r / rings
        for s in range(segments):
            ph = 2 * np.pi * s / segments
            pos.append((np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)))
    pos.append((0.0, -1.0, 0.0))
    pos = np.asarray(pos, np.float32)
    tris = []
    ring = lambda r, s: 1 + (r - 1) * segments + s % segments
    for s in range(segments):
        tris.append((0, ring(1, s + 1), ring(1, s)))
    for r in range(1, rings - 1):
        for s in range(segments):
            a, b, c, d = ring(r, s), ring(r, s + 1), ring(r + 1, s), ring(r + 1, s + 1)
            tris += [(a, b, d), (a, d, c)]
    last = len(pos) - 1
    for s in range(segments):
        tris.append((last, ring(rings - 1, s), ring(rings - 1, s + 1)))
    return pos, np.asarray(tris, np.uint32)


def quantize_snorm16(pos):
    """Positions as normalized int16 in [-1, 1] of their bounding cube, and
    the node (translation, scale) that maps them back."""
    lo, hi = pos.min(0), pos.max(0)
    c = (lo + hi) / 2
    s = float((hi - lo).max() / 2) or 1.0
    q = np.clip(np.rint((pos - c) / s * 32767), -32767, 32767).astype(np.int16)
    return q, c.astype(np.float64), s


def rotation_quat(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.r_[axis * np.sin(angle / 2), np.cos(angle / 2)]


def trs_matrix(t, q, s):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(s, np.float64)[None, :]
    M[:3, 3] = t
    return M


def instanced_spheres(grid=8, spacing=2.5, rings=24, segments=32, seed=0, glb_path=None, gltf_path=None):
    """A grid x grid field of quantized spheres (KHR_mesh_quantization
    positions, uint16 indices) under a root node with a matrix, three
    materials, one primitive without a material; returns the builder."""
    rng = np.random.default_rng(seed)
    b = Builder()
    b.js["extensionsUsed"] = ["KHR_mesh_quantization"]
    b.js["extensionsRequired"] = ["KHR_mesh_quantization"]
    pos, tris = uv_sphere(rings, segments)
    q, c, s = quantize_snorm16(pos)
    pa = b.accessor(q, normalized=True, minmax=True)
    ia = b.accessor(tris.astype(np.uint16).reshape(-1, 1))
    mats = [b.material(rng.random(3).tolist() + [1.0]) for _ in range(3)]
    meshes = [b.mesh([{"attributes": {"POSITION": pa}, "indices": ia, "material": m}]) for m in mats]
    meshes.append(b.mesh([{"attributes": {"POSITION": pa}, "indices": ia}]))
    kids = []
    for i in range(grid):
        for j in range(grid):
            ang = rng.uniform(0, 2 * np.pi)
            kids.append(b.node(mesh=int(rng.integers(0, len(meshes))),
                               translation=[float((i - grid / 2) * spacing), 0.0, float((j - grid / 2) * spacing)],
                               rotation=rotation_quat([0, 1, 0], ang).tolist(),
                               scale=[s, s, s]))
    root_m = trs_matrix([0.0, -1.0, 0.0], rotation_quat([1, 0, 0], 0.1), [1.0, 1.0, 1.0])
    root = b.node(matrix=root_m.T.ravel().tolist(), children=kids)
    b.scene([root])
    return b
