"""Host library (libvkgv_host.so): meshlet builder, normal cones, scene
generator and World traversal, checked on the CPU against brute force and the
closed forms of assets.cpp / world.cpp."""
import ctypes as C

import numpy as np
import pytest

import oracle
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import Scene, build_meshlets

from helpers import icosphere


def meshlet_triangles(m):
    out = []
    for ml in m.meshlets:
        vo, to, vc, tc = int(ml["vertex_offset"]), int(ml["triangle_offset"]), int(ml["vertex_count"]), int(ml["triangle_count"])
        loc = m.triangle_indices[to:to + 3 * tc].reshape(-1, 3).astype(np.int64)
        assert loc.max() < vc
        out.append(m.vertex_indices[vo + loc])
    return out


@pytest.mark.parametrize("level", [1, 3, 4])
def test_meshlets_partition_the_mesh(level):
    """Limits of the reference (64 vertices, 124 triangles, assets.cpp:325-346);
    every input triangle appears exactly once (same winding); triangle
    offsets 4-byte aligned (assets.cpp:339)."""
    pos, idx = icosphere(level)
    m = build_meshlets(pos, idx)
    assert np.all(m.meshlets["vertex_count"] <= 64) and np.all(m.meshlets["triangle_count"] <= 124)
    assert np.all(m.meshlets["triangle_offset"] % 4 == 0)
    got = np.concatenate(meshlet_triangles(m))
    canon = lambda t: tuple(np.roll(t, -int(np.argmin(t))))
    want = sorted(canon(t) for t in idx.reshape(-1, 3))
    assert sorted(canon(t) for t in got) == want


def test_meshlet_aabb_closed_form():
    """Meshlet AABB = (min + max) / 2, (max - min) / 2 of its vertex positions
    (assets.cpp:349-372)."""
    pos, idx = icosphere(3)
    pos = pos * np.float32(3.5) + np.float32([1, -2, 0.5])
    m = build_meshlets(pos, idx)
    for ml in m.meshlets:
        vo, vc = int(ml["vertex_offset"]), int(ml["vertex_count"])
        p = pos[m.vertex_indices[vo:vo + vc]]
        mn, mx = p.min(0), p.max(0)
        assert np.allclose(ml["aabb_center"], (mn + mx) * np.float32(0.5), atol=1e-6)
        assert np.allclose(ml["aabb_extents"], (mx - mn) * np.float32(0.5), atol=1e-6)


def test_cones_are_conservative():
    """A normal cone never culls a meshlet that has a triangle facing the eye
    (brute force over random eyes and transforms)."""
    pos, idx = icosphere(4)
    rng = np.random.default_rng(21)
    m = build_meshlets(pos, idx)
    tris = meshlet_triangles(m)
    culled_any = 0
    for _ in range(40):
        eye = rng.normal(size=3) * rng.uniform(1.2, 6)
        ang = rng.uniform(0, 2 * np.pi)
        R = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]])
        M = np.eye(4)
        M[:3, :3] = R * 1.7
        M[:3, 3] = rng.uniform(-1, 1, 3)
        mcol = np.ascontiguousarray(M.T.ravel(), np.float32)
        eye32 = np.asarray(eye, np.float32)
        for k, t in enumerate(tris):
            cone = N.MeshletCone.from_buffer_copy(m.cones[k].tobytes())
            if not oracle.lib().ref_cone_culled(C.addressof(cone), mcol.ctypes.data, eye32.ctypes.data):
                continue
            culled_any += 1
            w = (M @ np.c_[pos[t.ravel()].astype(np.float64), np.ones(t.size)].T).T[:, :3].reshape(-1, 3, 3)
            n = np.cross(w[:, 1] - w[:, 0], w[:, 2] - w[:, 0])
            assert np.all((n * (eye - w[:, 0])).sum(1) <= 1e-9), k
    assert culled_any > 100


def test_scene_generator_is_deterministic_and_sized():
    a = Scene(1, 1).arrays()
    b = Scene(1, 1).arrays()
    assert np.array_equal(a.draws, b.draws) and np.array_equal(a.transforms, b.transforms)
    assert all(np.array_equal(x.vertices, y.vertices) for x, y in zip(a.meshes, b.meshes))
    assert Scene(1, 1).triangle_count == 256 * 5120             # 256 subdivision-4 icospheres
    c2 = Scene(2, 2)
    assert 10_000_000 < c2.triangle_count < 11_000_000           # the ~10.5M-triangle city
    assert Scene(1, 2).arrays().transforms.tobytes() != a.transforms.tobytes()
    v = a.views[0]
    assert (v.width, v.height) == (1280, 720)
    assert len(Scene(3, 3).arrays().views) == 64


def test_world_build_depth_first_order():
    """World::iterateNode order (world.cpp:187-228): depth first, children in
    order; transform = parent * T * R * S; one draw per (primitive, meshlet)."""
    L = N.host()
    nodes = (N.Node * 5)()
    # 0 (mesh 0) -> [1 (no mesh) -> [3 (mesh 1)], 2 (mesh 0)];  4 (mesh 1) root
    spec = [(0, 0, 2), (-1, 2, 1), (0, 0, 0), (1, 0, 0), (1, 0, 0)]
    for i, (mesh, fc, cc) in enumerate(spec):
        nodes[i].mesh, nodes[i].first_child, nodes[i].child_count = mesh, fc, cc
        nodes[i].translation[:] = [i, 2 * i, 0]
        nodes[i].rotation[:] = [0, 0, 0, 1]
        nodes[i].scale[:] = [1 + i, 1, 1]
    children = np.array([1, 2, 3], np.uint32)
    roots = np.array([0, 4], np.uint32)
    mesh_prims = np.array([0, 1, 1, 2], np.uint32)      # mesh 0: prim 0; mesh 1: prims 1, 2
    prim_list = np.array([0, 1, 2], np.uint32)
    meshlet_counts = np.array([2, 1, 3], np.uint32)
    nd, nt = C.c_uint32(0), C.c_uint32(0)
    args = [nodes, 5, children.ctypes.data, roots.ctypes.data, 2, mesh_prims.ctypes.data, prim_list.ctypes.data,
            meshlet_counts.ctypes.data]
    assert L.vkgv_world_build(*args, None, C.byref(nd), None, C.byref(nt)) == 0
    draws = np.zeros(nd.value, dtype=[("p", "<u4"), ("m", "<u4"), ("t", "<u4")])
    tr = np.zeros((nt.value, 16), np.float32)
    assert L.vkgv_world_build(*args, draws.ctypes.data, C.byref(nd), tr.ctypes.data, C.byref(nt)) == 0
    # visit order of mesh nodes: 0, 3, 2, 4 -> transforms 0..3
    want = [(0, 0, 0), (0, 1, 0),                                   # node 0: prim 0 x 2 meshlets
            (1, 0, 1), (2, 0, 1), (2, 1, 1), (2, 2, 1),             # node 3: prims 1, 2
            (0, 0, 2), (0, 1, 2),                                   # node 2
            (1, 0, 3), (2, 0, 3), (2, 1, 3), (2, 2, 3)]             # node 4
    assert [tuple(int(x) for x in d) for d in draws] == want
    T = lambda i: np.array([[1 + i, 0, 0, i], [0, 1, 0, 2 * i], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    for k, chain in enumerate([[0], [0, 1, 3], [0, 2], [4]]):
        M = np.eye(4)
        for n in chain:
            M = M @ T(n)
        assert np.allclose(tr[k].reshape(4, 4).T, M), k


class _FakeRenderer:
    """Stands in for a Renderer: a frame's duration depends on the tuning."""

    def __init__(self, cost):
        self.t, self.cost = (16, 3, 16), cost

    def set_tuning(self, *t):
        self.t = tuple(t)

    def get_tuning(self):
        return self.t

    def sync(self):
        pass

    def frame(self):
        import time
        time.sleep(self.cost(self.t))


def test_autotune_picks_fastest_valid_candidate(tmp_path, monkeypatch):
    from vk_gltf_viewer_amd.autotune import autotune
    monkeypatch.setenv("VKGV_TUNE_CACHE", str(tmp_path / "tune.json"))
    cand = {"small_max_extent": (4, 16), "raster_blocks_per_cu": (2, 3), "tile_item_tris": (16,)}
    # small extent 4 is the fastest but "overflows"
    r = _FakeRenderer(lambda t: 0.0005 if t[0] == 4 else (0.002 if t[1] == 2 else 0.001))
    best, times = autotune(r, r.frame, "k", candidates=cand, warmup=0, frames=2, use_cache=True,
                           valid=lambda: r.t[0] != 4, search="grid")
    assert best == (16, 3, 16) and r.t == best
    assert times[(4, 2, 16)] == float("inf") and times[(4, 3, 16)] == float("inf")
    # cached: applied without timing
    r.set_tuning(4, 2, 16)
    best2, times2 = autotune(r, r.frame, "k", candidates=cand, use_cache=True)
    assert best2 == best and r.t == best and times2 == {}
    # every candidate rejected: the tuning is left as it was
    r.set_tuning(32, 2, 8)
    best3, _ = autotune(r, r.frame, "k2", candidates=cand, warmup=0, frames=1, use_cache=False, valid=lambda: False)
    assert best3 == (32, 2, 8) and r.t == (32, 2, 8)


def test_autotune_coordinate_search_and_stale_cache(tmp_path, monkeypatch):
    """The coordinate search finds the optimum of a separable cost with fewer
    timings than the grid; cache entries of the wrong length, or whose tuning
    no longer passes `valid`, are searched again; the key names the device and
    the knob count."""
    import json

    from vk_gltf_viewer_amd.autotune import KNOBS, autotune, cache_key
    path = tmp_path / "tune.json"
    monkeypatch.setenv("VKGV_TUNE_CACHE", str(path))
    cand = {"small_max_extent": (4, 8, 16, 32), "raster_blocks_per_cu": (2, 3), "tile_item_tris": (8, 16, 32)}
    best_t = (8, 2, 32)
    cost = lambda t: 0.0002 * (1 + sum(a != b for a, b in zip(t, best_t)))
    r = _FakeRenderer(cost)
    best, times = autotune(r, r.frame, "sep", candidates=cand, warmup=0, frames=1)
    assert best == best_t and r.t == best_t
    assert len(times) < 4 * 2 * 3
    key = cache_key(r, "sep")
    assert key.startswith("sep|") and f"knobs-{len(KNOBS)}" in key
    data = json.loads(path.read_text())
    assert data[key]["tuning"] == list(best_t)
    # an entry of the wrong length (an older build's knob set) is not applied
    data[key]["tuning"] = [4, 3]
    path.write_text(json.dumps(data))
    r.set_tuning(16, 3, 16)
    best2, times2 = autotune(r, r.frame, "sep", candidates=cand, warmup=0, frames=1)
    assert best2 == best_t and times2
    # a cached tuning that fails `valid` now is searched again
    best3, times3 = autotune(r, r.frame, "sep", candidates=cand, warmup=0, frames=1, valid=lambda: r.t[0] != 8)
    assert best3[0] != 8 and times3
