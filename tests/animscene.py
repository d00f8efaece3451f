"""Synthetic node hierarchies with animations (every interpolation and path)
for the transform-update tests."""
import numpy as np

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import ANIM_CHANNEL_DTYPE, ANIM_SAMPLER_DTYPE, NODE_DTYPE, Hierarchy


def unit_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def random_hierarchy(seed, nodes=200, roots=3, mesh_frac=0.6, animated_frac=0.5, keys=(2, 9)):
    """A forest of `roots` trees over `nodes` nodes (children in index order),
    static TRS, and samplers / channels: LINEAR, STEP and CUBICSPLINE on
    translation, rotation and scale, some nodes driven twice on one path."""
    rng = np.random.default_rng(seed)
    parent = np.full(nodes, -1)
    for i in range(roots, nodes):
        parent[i] = rng.integers(0, i)
    kids = [[] for _ in range(nodes)]
    for i in range(nodes):
        if parent[i] >= 0:
            kids[parent[i]].append(i)
    nd = np.zeros(nodes, NODE_DTYPE)
    children = []
    for i in range(nodes):
        nd[i]["mesh"] = int(rng.integers(0, 4)) if rng.random() < mesh_frac else -1
        nd[i]["first_child"], nd[i]["child_count"] = len(children), len(kids[i])
        children += kids[i]
        nd[i]["translation"] = rng.normal(0, 2, 3)
        nd[i]["rotation"] = unit_quats(rng, 1)[0]
        nd[i]["scale"] = rng.uniform(0.5, 1.5, 3)
    samplers, channels, times, values = [], [], [], []
    nt = nv = 0
    for i in range(nodes):
        if rng.random() >= animated_frac:
            continue
        for _ in range(int(rng.integers(1, 4))):
            path = int(rng.integers(0, 3))
            interp = int(rng.integers(0, 3))
            k = int(rng.integers(*keys))
            t = rng.uniform(0, 4, k).astype(np.float32)
            t[0] = rng.uniform(0, 0.5)
            t = np.sort(t)                                 # glTF keyframe times ascend
            comps = 4 if path == N.PATH_ROTATION else 3
            outs = k * (3 if interp == N.INTERP_CUBICSPLINE else 1)
            if path == N.PATH_ROTATION:
                v = unit_quats(rng, outs)
                if interp == N.INTERP_CUBICSPLINE:
                    v = v.reshape(k, 3, 4)
                    v[:, [0, 2]] *= 0.3                # tangents
                    v = v.reshape(-1, 4)
            elif path == N.PATH_SCALE:
                v = rng.uniform(0.5, 2.0, (outs, 3)).astype(np.float32)
            else:
                v = rng.normal(0, 2, (outs, 3)).astype(np.float32)
            samplers.append((nt, k, nv, interp))
            channels.append((i, len(samplers) - 1, path, 0))
            times.append(t)
            values.append(v.ravel())
            nt += k
            nv += v.size
    order = rng.permutation(len(channels))          # channels interleaved across nodes
    h = Hierarchy(nd, np.asarray(children, np.uint32), np.arange(roots, dtype=np.uint32),
                  np.array(samplers, np.uint32).reshape(-1, 4).view(ANIM_SAMPLER_DTYPE).ravel(),
                  np.concatenate(times) if times else np.zeros(0, np.float32),
                  np.concatenate(values) if values else np.zeros(0, np.float32),
                  np.array(channels, np.uint32)[order].reshape(-1, 4).view(ANIM_CHANNEL_DTYPE).ravel())
    return h
