"""Device transform update (vkgv_set_hierarchy / vkgv_update_transforms) and
motion vectors (VKGV_MOTION_VECTORS) against the oracle on the same inputs:
matrices bit-exact for random animated forests at several times; a frame
after the device update equals the frame after uploading the same matrices
from the host; motion vectors bit-exact in single-view and batched frames."""
import numpy as np
import pytest

import oracle
from animscene import random_hierarchy
from gltfgen import instanced_spheres
from test_gpu_parity import compare_color, compare_keys
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd import gltf
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import (ANIM_CHANNEL_DTYPE, ANIM_SAMPLER_DTYPE, Hierarchy, camera_advance,
                                      camera_from_view, view_desc)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


@pytest.mark.parametrize("seed,nodes", [(1, 300), (2, 5000), (3, 40)])
def test_update_matches_oracle(rend, seed, nodes):
    h = random_hierarchy(seed, nodes=nodes, roots=1 + seed, animated_frac=0.6)
    print(f"hierarchy seed {seed}: {nodes} nodes, {len(h.channels)} channels", flush=True)
    rend.set_hierarchy(h)
    count = int((h.nodes["mesh"] >= 0).sum())      # every node is reachable from the roots
    for time in (0.0, 0.37, 1.5, 3.99, 11.25):
        rend.update_transforms(time)
        got = rend.read_transforms(count)
        want = oracle.update_transforms(h, time)
        assert want.shape == got.shape
        same = np.array_equal(got.view(np.uint32), want.view(np.uint32))
        print(f"  t={time}: exact={same}, max |diff| {np.abs(got - want).max():.3g}", flush=True)
        assert same


def animated_spheres(tmp_path, seed=4):
    p = tmp_path / "spheres.glb"
    instanced_spheres(grid=10, spacing=2.2, seed=seed).write_glb(p)
    sc = gltf.load_gltf(str(p))
    h = sc.hierarchy
    rng = np.random.default_rng(seed)
    # drive the root (rotation) and every third child (translation, scale) with keyframes
    samplers, channels, times, values = [], [], [], []
    nt = nv = 0

    def add(node, path, interp, t, v):
        nonlocal nt, nv
        samplers.append((nt, len(t), nv, interp))
        channels.append((node, len(samplers) - 1, path, 0))
        times.append(np.asarray(t, np.float32))
        values.append(np.asarray(v, np.float32).ravel())
        nt += len(t)
        nv += values[-1].size
    root = int(h.roots[0])
    q = lambda a: [0.0, np.sin(a / 2), 0.0, np.cos(a / 2)]
    add(root, N.PATH_ROTATION, N.INTERP_LINEAR, [0.0, 2.0, 4.0], [q(0.0), q(0.4), q(0.0)])
    kids = h.children[int(h.nodes[root]["first_child"]):][:int(h.nodes[root]["child_count"])]
    for k in kids[::3]:
        t0 = h.nodes[int(k)]["translation"]
        add(int(k), N.PATH_TRANSLATION, N.INTERP_LINEAR, [0.0, 1.0, 3.0],
            [t0, t0 + rng.normal(0, 0.5, 3), t0])
        add(int(k), N.PATH_SCALE, N.INTERP_STEP, [0.0, 1.5, 3.0], rng.uniform(0.1, 0.3, (3, 3)))
    h2 = Hierarchy(h.nodes, h.children, h.roots,
                   np.array(samplers, np.uint32).reshape(-1, 4).view(ANIM_SAMPLER_DTYPE).ravel(),
                   np.concatenate(times), np.concatenate(values),
                   np.array(channels, np.uint32).reshape(-1, 4).view(ANIM_CHANNEL_DTYPE).ravel())
    return sc, h2


def test_frame_after_device_update_equals_host_upload(rend, tmp_path):
    sc, h = animated_spheres(tmp_path)
    rend.load(sc)
    W, H = 480, 320
    v = rend.create_view(W, H)
    cam = camera_from_view(view_desc(W, H, (0.0, 8.0, 16.0), (0.0, -1.0, 0.0), fovy_deg=55.0))
    try:
        rs_flags = N.CULL_FRUSTUM | N.RESOLVE_BARY
        for time in (0.6, 2.3):
            print(f"animated frame t={time}", flush=True)
            rend.set_hierarchy(h)
            rend.update_transforms(time)
            v.render(cam, rs_flags)
            k1, c1 = v.keys().copy(), v.color().copy()
            T = oracle.update_transforms(h, time)
            rend.set_transforms(T)
            v.render(cam, rs_flags)
            assert np.array_equal(v.keys(), k1) and np.array_equal(v.color(), c1)
            sc.transforms = T
            rs = oracle.Scene(sc)
            _, _, total = oracle.pyramid_info(W, H)
            rk, rc, _ = oracle.frame_reference(rs, cam, W, H, rs_flags, np.zeros(total, np.float32))
            compare_keys(k1, rk, f"animated t={time}")
            compare_color(c1, rc, f"animated t={time}")
            assert np.array_equal(k1, rk) and np.array_equal(c1, rc)
    finally:
        v.close()


def test_motion_vectors_match_oracle(rend, tmp_path):
    sc, h = animated_spheres(tmp_path, seed=6)
    rend.load(sc)
    rend.set_hierarchy(h)
    specs = [(400, 300, (0.0, 7.0, 15.0)), (256, 384, (5.0, 4.0, 12.0))]
    views = [rend.create_view(W, H) for W, H, _ in specs]
    flags = N.CULL_FRUSTUM | N.MOTION_VECTORS
    try:
        t0, t1 = 0.5, 0.8
        cams = []
        for (W, H, e) in specs:
            c0 = camera_from_view(view_desc(W, H, e, (0.0, -1.0, 0.0), fovy_deg=50.0))
            cams.append(camera_advance(c0, view_desc(W, H, (e[0] + 0.3, e[1], e[2] - 0.2), (0.0, -1.0, 0.0),
                                                     fovy_deg=50.0)))
        rend.update_transforms(t0)
        rend.update_transforms(t1)
        prev, cur = oracle.update_transforms(h, t0), oracle.update_transforms(h, t1)
        sc.transforms = cur
        rs = oracle.Scene(sc)
        got = []
        for v, c, (W, H, _) in zip(views, cams, specs):
            print(f"motion vectors {W}x{H}", flush=True)
            v.render(c, flags)
            keys = v.keys()
            mv = v.motion()
            want = oracle.motion_vectors(rs, c, prev, keys, W, H)
            covered = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
            assert covered.mean() > 0.2
            assert np.abs(want[covered]).max() > 1e-3                  # the frame moves
            same = np.array_equal(mv.view(np.uint32), want.view(np.uint32))
            print(f"  exact={same}, max |diff| {np.abs(mv - want).max():.3g}", flush=True)
            assert same
            got.append(mv.copy())
        # batched: the same vectors
        batch = rend.create_batch(views)
        try:
            batch.render(cams, flags)
            rend.sync()
            for v, g in zip(views, got):
                assert np.array_equal(v.motion(), g)
        finally:
            batch.close()
    finally:
        for v in views:
            v.close()


def test_state_errors(rend, tmp_path):
    r = Renderer()
    try:
        with pytest.raises(VkgvError, match="hierarchy"):
            r.update_transforms(0.0)
        h = random_hierarchy(9, nodes=30)
        r.set_hierarchy(h)
        r.update_transforms(1.0)
        n = int((h.nodes["mesh"] >= 0).sum())
        r.set_transforms(np.tile(np.eye(4, dtype=np.float32).ravel(), (n + 1, 1)))
        with pytest.raises(VkgvError, match="replaced"):
            r.update_transforms(1.0)
        bad = random_hierarchy(9, nodes=30)
        bad.channels = bad.channels.copy()
        bad.channels[0]["sampler"] = 10 ** 6
        with pytest.raises(VkgvError, match="channel"):
            r.set_hierarchy(bad)
        r.sync()
    finally:
        r.close()


def test_graph_replays_with_device_updates(rend, tmp_path):
    """A captured batch frame (VKGV_GRAPH) with the textured resolve and motion
    vectors, replayed after each device transform update: every replay equals
    the eager batch frame of the same state (colour, keys, motion vectors)."""
    from texscene import textured_spheres
    p = tmp_path / "ts.glb"
    textured_spheres(grid=6, seed=2).write_glb(p)
    sc = gltf.load_gltf(str(p))
    h = sc.hierarchy
    # spin the scene's roots about y so that the transforms change between updates
    q = lambda a: [0.0, float(np.sin(a / 2)), 0.0, float(np.cos(a / 2))]
    nr = len(h.roots)
    sp = np.array([(0, 3, 0, N.INTERP_LINEAR)], np.uint32).view(ANIM_SAMPLER_DTYPE).ravel()
    times = np.array([0.0, 1.0, 2.0], np.float32)
    values = np.array([q(0.0), q(0.5), q(1.0)], np.float32).ravel()
    ch = np.zeros(nr, ANIM_CHANNEL_DTYPE)
    ch["node"], ch["sampler"], ch["path"] = h.roots, 0, N.PATH_ROTATION
    h2 = Hierarchy(h.nodes, h.children, h.roots, sp, times, values, ch)
    rend.load(sc)
    rend.set_hierarchy(h2)
    specs = [(320, 240, (0.0, 6.0, 12.0)), (200, 160, (5.0, 3.0, 9.0))]
    views = [rend.create_view(W, H) for W, H, _ in specs]
    cams = [camera_from_view(view_desc(W, H, e, (0.0, 0.0, 0.0), fovy_deg=50.0)) for W, H, e in specs]
    flags = N.CULL_FRUSTUM | N.RESOLVE_TEXTURED | N.MOTION_VECTORS
    batch = rend.create_batch(views)
    try:
        t_prev = 0.0                                  # set_hierarchy leaves the time-0 matrices
        for k, t in enumerate((0.2, 0.45, 0.7, 1.3)):
            print(f"graph replay {k} at t={t}", flush=True)
            rend.update_transforms(t)                 # previous = t_prev, current = t
            batch.render(cams, flags | N.GRAPH)
            rend.sync()
            got = [(v.keys().copy(), v.color().copy(), v.motion().copy()) for v in views]
            # the same state rebuilt eagerly: previous = t_prev, current = t
            rend.update_transforms(t_prev)
            rend.update_transforms(t)
            batch.render(cams, flags)
            rend.sync()
            for v, (gk, gc, gm) in zip(views, got):
                assert np.array_equal(v.keys(), gk)
                assert np.array_equal(v.color(), gc)
                assert np.array_equal(v.motion().view(np.uint32), gm.view(np.uint32))
            assert np.abs(got[0][2]).max() > 0                        # the scene moves
            t_prev = t
    finally:
        batch.close()
        for v in views:
            v.close()
