"""Full queues.  A triangle that finds the big-triangle queue, the clip queue
or the chunk list full is rasterized by the direct path (the wave sweeps its
bounding box), and one that finds a screen tile's list full is swept over
that tile, so the frame does not change: keys, colours and visible lists are
bit-identical to those of a frame with room to spare, and to the oracle,
while the overflow bits say which queue was full.

The scene: a field of spheres over a tiled ground plane that runs under and
behind the camera, with the near plane cutting the nearest spheres and the
ground (clipping), large ground tiles (big triangles), and, with the
small-triangle extent at one pixel, more big triangles than the chunk list
has chunks (640x480: 2^15)."""
import numpy as np
import pytest

import oracle
from gltfgen import instanced_spheres
from test_gpu_parity import compare_color, compare_keys
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd import gltf
from vk_gltf_viewer_amd.renderer import Renderer
from vk_gltf_viewer_amd.scene import camera_from_view, view_desc

pytestmark = pytest.mark.gpu

W, H = 640, 480
DEFAULT_TUNING = (16, 3, 8, 8192)
ONE_PIXEL = (1, 3, 8, 8192)          # every triangle of a bounding box >= 1 px goes to the queues
FLAGS = [N.MODE_REFERENCE, N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY]


def crowded_scene(tmp_path):
    b = instanced_spheres(grid=12, spacing=2.2, rings=40, segments=64, seed=3)
    n, ext = 24, 40.0
    xs = np.linspace(-ext, ext, n + 1, dtype=np.float32)
    gx, gz = np.meshgrid(xs, xs)
    pos = np.stack([gx.ravel(), np.full(gx.size, -2.2, np.float32), gz.ravel()], 1).astype(np.float32)
    v = (np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]).ravel()
    tris = np.concatenate([np.stack([v, v + n + 1, v + 1], 1), np.stack([v + 1, v + n + 1, v + n + 2], 1)])
    pa = b.accessor(pos, minmax=True)
    ia = b.accessor(tris.astype(np.uint32).reshape(-1, 1))
    m = b.material([0.35, 0.5, 0.3, 1.0], double_sided=True)
    ground = b.node(mesh=b.mesh([{"attributes": {"POSITION": pa}, "indices": ia, "material": m}]))
    b.js["scenes"][0]["nodes"].append(ground)
    p = tmp_path / "crowded.glb"
    b.write_glb(p)
    return gltf.load_gltf(str(p))


def camera():
    return camera_from_view(view_desc(W, H, (1.1, -1.0, 1.1), (8.0, -2.0, 6.0), fovy_deg=70.0, znear=0.6))


def frames(rend, view, cam, flags):
    """Two frames from a cleared history and an all-far pyramid, with stats."""
    out = []
    view.reset_history()
    view.write_pyramid(np.zeros(view.pyramid_texels, np.float32))
    for _ in range(2):
        view.render(cam, flags)
        st = view.stats()
        out.append((view.keys().copy(), view.color().copy(), np.sort(view.visible(0)), np.sort(view.visible(1)),
                    st.overflow, st.clipped_triangles, st.big_triangles))
    return out


@pytest.fixture(scope="module")
def crowded(tmp_path_factory):
    return crowded_scene(tmp_path_factory.mktemp("overflow"))


def run(sc, tuning, **caps):
    r = Renderer(**caps)
    try:
        r.load(sc)
        r.set_tuning(*tuning)
        view = r.create_view(W, H)
        try:
            return {f: frames(r, view, camera(), f) for f in FLAGS}
        finally:
            view.close()
    finally:
        r.close()


def test_roomy_frame_matches_oracle(crowded):
    print("roomy queues, default tuning", flush=True)
    got = run(crowded, DEFAULT_TUNING)
    fr = got[N.MODE_REFERENCE]
    assert all(f[4] == 0 for f in fr)                       # no queue was full
    assert fr[0][5] > 0 and fr[0][6] > 64                   # the scene clips and bins
    rs = oracle.Scene(crowded)
    _, _, total = oracle.pyramid_info(W, H)
    pyr = np.zeros(total, np.float32)
    for k in range(2):
        rk, rc, rv = oracle.frame_reference(rs, camera(), W, H, N.MODE_REFERENCE, pyr)
        compare_keys(fr[k][0], rk, f"roomy frame {k}")
        compare_color(fr[k][1], rc, f"roomy frame {k}")
        assert np.array_equal(fr[k][0], rk) and np.array_equal(fr[k][1], rc)
        assert np.array_equal(fr[k][2], np.sort(rv))


@pytest.mark.parametrize("tuning,caps,bits", [
    (DEFAULT_TUNING, dict(big_triangle_capacity=16, clip_triangle_capacity=1), 0b011),
    (ONE_PIXEL, {}, 0b100),
    (ONE_PIXEL, dict(big_triangle_capacity=40000, clip_triangle_capacity=2), 0b110),
    (ONE_PIXEL, dict(big_triangle_capacity=8, clip_triangle_capacity=1), 0b011),
    (DEFAULT_TUNING, dict(tile_triangle_capacity=1), 0b1000),
    (ONE_PIXEL, dict(tile_triangle_capacity=3), 0b1000),
])
def test_full_queues_give_the_same_frame(crowded, tuning, caps, bits):
    print(f"tuning {tuning}, capacities {caps}", flush=True)
    ref = run(crowded, DEFAULT_TUNING)
    got = run(crowded, tuning, **caps)
    for f in FLAGS:
        for k, (a, b) in enumerate(zip(ref[f], got[f])):
            print(f"  flags {f:#x} frame {k}: overflow {b[4]:#x}, clipped {b[5]}, big {b[6]}", flush=True)
            for i in range(4):
                assert np.array_equal(a[i], b[i]), (f, k, i)
        assert any((b[4] & bits) == bits for b in got[f]), (f, [b[4] for b in got[f]])


def test_full_queues_in_a_batch(crowded):
    """Two views in one batch frame, with queues too small for either: each
    view's image equals its single-view frame with room to spare."""
    specs = [(W, H, camera()),
             (331, 217, camera_from_view(view_desc(331, 217, (0.0, -0.6, -3.0), (4.0, -1.8, 5.0), fovy_deg=80.0,
                                                   znear=0.8)))]   # odd size: sweeps and tiles cross the edges
    ref = []
    r = Renderer()
    try:
        r.load(crowded)
        for w, h, cam in specs:
            v = r.create_view(w, h)
            v.render(cam, N.MODE_REFERENCE)
            ref.append((v.keys().copy(), v.color().copy()))
            assert v.stats().overflow == 0
            v.close()
    finally:
        r.close()
    r = Renderer(big_triangle_capacity=8, clip_triangle_capacity=1, tile_triangle_capacity=1)
    try:
        r.load(crowded)
        views = [r.create_view(w, h) for w, h, _ in specs]
        batch = r.create_batch(views)
        try:
            batch.render([c for _, _, c in specs], N.MODE_REFERENCE)
            r.sync()
            assert batch.stats().overflow & 0b11
            for v, (k, c) in zip(views, ref):
                assert np.array_equal(v.keys(), k) and np.array_equal(v.color(), c)
            print("batch with full queues: both views identical to their roomy single-view frames", flush=True)
            # the same frame captured as a graph and replayed, with eager frames in between
            flags = N.CULL_FRUSTUM | N.RESOLVE_BARY
            batch.render([c for _, _, c in specs], flags)
            r.sync()
            eager = [(v.keys().copy(), v.color().copy()) for v in views]
            for rep in range(3):
                batch.render([c for _, _, c in specs], flags | N.GRAPH)
                r.sync()
                for v, (k, c) in zip(views, eager):
                    assert np.array_equal(v.keys(), k) and np.array_equal(v.color(), c), rep
                batch.render([c for _, _, c in specs], flags)
                r.sync()
            print("graph replays with full queues identical to the eager frame", flush=True)
        finally:
            batch.close()
            for v in views:
                v.close()
    finally:
        r.close()
