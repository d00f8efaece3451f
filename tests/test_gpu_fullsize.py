"""The benchmarked workloads at their full size.

- The occlusion-heavy 4K city (10.5M triangles) with the camera moving down
  the street: five two-pass frames against the oracle (early and late sets,
  pyramid, keys, colour), so the late pass and the history change every frame.
- The literal city of 4096 level-5 spheres (83.9M triangles) at 4K: the
  reference frame (previous-frame Hi-Z) against the oracle, frame 0 with an
  all-far pyramid and frame 1 against the pyramid frame 0 built.
- All 64 views of the batched configuration: batch frames equal single-view
  frames bit for bit, over a static and a moved frame.
- The uneven stress configuration at full size (200 meshes, 15.2M draws,
  16 views up to 4K): captured batch frames equal single-view frames.
"""
import functools

import numpy as np
import pytest

import oracle
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer
from vk_gltf_viewer_amd.scene import Scene, camera_advance, camera_from_view, view_desc

pytestmark = pytest.mark.gpu

TWO_PASS = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY


@functools.lru_cache(maxsize=2)
def scene(cfg, seed, variant=0):
    return Scene(cfg, seed, variant).arrays()


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


def test_c2_moving_camera_two_pass_matches_oracle(rend):
    sc = scene(2, 2)
    rend.load(sc)
    vd = N.ViewDesc.from_buffer_copy(sc.views[0])
    W, H = vd.width, vd.height
    view = rend.create_view(W, H)
    try:
        rs = oracle.Scene(sc)
        hist = np.zeros(len(sc.draws), np.uint8)
        cam = camera_from_view(vd)
        lates = []
        for frame in range(5):
            if frame:
                vd.eye[0] += 1.5
                vd.target[0] += 1.5
                cam = camera_advance(cam, vd)
            print(f"case: c2 4K moving frame {frame}, eye x {vd.eye[0]:.1f}", flush=True)
            rk, rc, rp, re, rl = oracle.frame_two_pass(rs, cam, W, H, TWO_PASS, hist)
            view.render(cam, TWO_PASS)
            assert np.array_equal(view.visible(0), re), frame
            assert np.array_equal(view.visible(1), rl), frame
            assert np.array_equal(view.pyramid().view(np.uint32), rp.view(np.uint32)), frame
            assert np.array_equal(view.keys(), rk), frame
            assert np.array_equal(view.color(), rc), frame
            lates.append(len(rl))
            print(f"  early {len(re)} late {len(rl)}: sets, pyramid, keys and colour bit-exact", flush=True)
        # frame 0 is cold (everything in the late pass) and frame 1's early pass
        # draws all of frame 0's late list; from frame 2 on, the motion uncovers
        # draws that only the late pass finds
        assert all(n > 0 for n in lates[2:]), "the moving camera must uncover draws in the late pass"
    finally:
        view.close()


def test_c2_literal_reference_frames_match_oracle(rend):
    sc = scene(2, 2, 1)
    tris = sum(int(m.meshlets["triangle_count"].astype(np.int64).sum()) for m in sc.meshes[1:])
    print(f"case: literal city, {len(sc.meshes) - 1} spheres, {len(sc.draws)} draws, >= {tris} sphere triangles",
          flush=True)
    assert tris == 4096 * 20480
    rend.load(sc)
    vd = sc.views[0]
    W, H = vd.width, vd.height
    cam = camera_from_view(vd)
    view = rend.create_view(W, H)
    try:
        rs = oracle.Scene(sc)
        _, _, total = oracle.pyramid_info(W, H)
        pyr = np.zeros(total, np.float32)
        for frame in range(2):
            rk, rc, rv = oracle.frame_reference(rs, cam, W, H, N.MODE_REFERENCE, pyr)
            view.render(cam, N.MODE_REFERENCE)
            assert np.array_equal(view.visible(0), rv), frame
            assert np.array_equal(view.keys(), rk), frame
            assert np.array_equal(view.color(), rc), frame
            assert np.array_equal(view.pyramid().view(np.uint32), pyr.view(np.uint32)), frame
            print(f"  frame {frame}: {len(rv)} visible draws; keys, colour and pyramid bit-exact", flush=True)
    finally:
        view.close()


def moved(v, f):
    """The view with its eye moved a little in frame f (perspective views only)."""
    if v.orthographic or f == 0:
        return v
    return view_desc(v.width, v.height, (v.eye[0] + 0.5 * f, v.eye[1] + 0.2 * f, v.eye[2] - 0.3 * f), tuple(v.target),
                     fovy_deg=float(np.degrees(v.fovy)), up=tuple(v.up), znear=v.znear, zfar=v.zfar)


def _batch_equals_single(rend, vds, frames, flags, label):
    bviews = [rend.create_view(v.width, v.height) for v in vds]
    sviews = [rend.create_view(v.width, v.height) for v in vds]
    batch = rend.create_batch(bviews)
    try:
        cams = [None] * len(vds)
        for f in range(frames):
            cur = [moved(v, f) for v in vds]
            cams = [camera_from_view(v) if c is None else camera_advance(c, v) for v, c in zip(cur, cams)]
            batch.render(cams, flags)
            for sv, cam in zip(sviews, cams):
                sv.render(cam, TWO_PASS)
            drawn = 0
            for i, (bv, sv) in enumerate(zip(bviews, sviews)):
                kb = bv.keys()
                assert np.array_equal(kb, sv.keys()), (label, f, i)
                assert np.array_equal(bv.color(), sv.color()), (label, f, i)
                assert np.array_equal(bv.visible(0), sv.visible(0)), (label, f, i)
                assert np.array_equal(bv.visible(1), sv.visible(1)), (label, f, i)
                drawn += int((kb != np.uint64(N.KEY_CLEAR)).any())
            print(f"{label} frame {f}: {len(vds)} views bit-identical to single-view frames "
                  f"({drawn} with something drawn)", flush=True)
            assert drawn >= len(vds) // 2
        assert batch.stats().overflow == 0
    finally:
        batch.close()
        for v in bviews + sviews:
            v.close()


def test_c3_all_64_views_batch_equals_single_views(rend):
    sc = scene(3, 3)
    assert len(sc.views) == 64
    rend.load(sc)
    print("case: configuration 3, all 64 views", flush=True)
    _batch_equals_single(rend, sc.views, 2, TWO_PASS | N.GRAPH, "c3")


def test_c5_full_size_batch_equals_single_views(rend):
    sc = scene(5, 5, 0)
    print(f"case: configuration 5 variant 0: {len(sc.meshes)} meshes, {len(sc.draws)} draws", flush=True)
    assert len(sc.meshes) == 200 and len(sc.draws) > 15_000_000 and len(sc.views) == 16
    rend.load(sc)
    _batch_equals_single(rend, sc.views, 2, TWO_PASS | N.GRAPH, "c5")
