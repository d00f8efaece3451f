"""Error paths of the C ABI on the device.

- Allocation failures, injected from the host (VKGV_FAULT_ALLOC_AFTER=n makes
  a context's allocations fail after the first n): every allocation of the
  setup (context, scene, textures, transforms, draws, view, first frame) and
  of a later scene upload is made to fail in turn.  Each failure is reported
  as VKGV_ERROR_OUT_OF_MEMORY, and a failed upload leaves the context refusing
  frames with VKGV_ERROR_STATE instead of rendering from freed buffers.
- The draw-count limit of the 25-bit draw index: 2^25 draws render (the last
  one visible, so its index fills all 25 bits of the visibility-buffer id),
  2^25 + 1 are refused with VKGV_ERROR_LIMIT.
- A frame with no draws (the reference skips its draw, application.cpp:848)
  gives a cleared visibility buffer, a zero colour image and empty lists.
No kernel runs in a failed state: every failure is caught on the host.
"""
import os

import numpy as np
import pytest

import oracle
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import DRAW_DTYPE, Scene, camera_from_view

pytestmark = pytest.mark.gpu

FLAGS = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY


def _renderer_failing_after(n):
    os.environ["VKGV_FAULT_ALLOC_AFTER"] = str(n)
    try:
        return Renderer(0)
    finally:
        del os.environ["VKGV_FAULT_ALLOC_AFTER"]


def _setup(r, sc, cam, w, h):
    r.load(sc)
    v = r.create_view(w, h)
    v.render(cam, FLAGS)
    r.sync()
    return v


def test_allocation_failures_are_reported_and_leave_a_consistent_state():
    sc = Scene(1, 1, 16).arrays()
    vd = sc.views[0]
    w, h = 160, 90
    vd.width, vd.height = w, h
    cam = camera_from_view(vd)
    # the setup with its k-th allocation failing, for every k until it succeeds
    n_setup = None
    for k in range(200):
        try:
            r = _renderer_failing_after(k)
        except VkgvError as e:
            assert e.code == N.ERROR_HIP          # vkgv_create reports a failed set-up as a HIP error
            continue
        try:
            _setup(r, sc, cam, w, h)
            n_setup = k
            break
        except VkgvError as e:
            assert e.code == N.ERROR_OUT_OF_MEMORY, (k, str(e))
            # whatever is set, no frame can be rendered from it (no view can be made either)
            with pytest.raises(VkgvError):
                r.create_view(w, h)
        finally:
            r.close()
    assert n_setup is not None and n_setup > 20
    print(f"setup takes {n_setup} device allocations; each failure was reported", flush=True)
    # a second scene upload with its j-th allocation failing: frames are refused afterwards
    for j in range(40):
        r = _renderer_failing_after(n_setup + j)
        try:
            v = _setup(r, sc, cam, w, h)
            try:
                r.upload_scene(sc)
            except VkgvError as e:
                assert e.code == N.ERROR_OUT_OF_MEMORY, (j, str(e))
                with pytest.raises(VkgvError) as e2:
                    v.render(cam, FLAGS)
                assert e2.value.code == N.ERROR_STATE, (j, str(e2.value))
                with pytest.raises(VkgvError) as e3:
                    r.set_draws(sc.draws)
                assert e3.value.code in (N.ERROR_STATE, N.ERROR_OUT_OF_MEMORY), j
                continue
            # the upload went through (its last allocations were within the budget)
            print(f"a scene upload takes {j} device allocations; each failure left frames refused", flush=True)
            break
        finally:
            r.close()
    else:
        raise AssertionError("the second upload never succeeded")


def test_draw_count_limit():
    """2^25 draws: all but the last reference a transform behind the camera;
    the last one (draw index 2^25 - 1) is in view and must own the pixels a
    one-draw scene of the same meshlet gives, with the full 25-bit index."""
    sc = Scene(1, 1, 1).arrays()
    vd = sc.views[0]
    w, h = 320, 180
    vd.width, vd.height = w, h
    cam = camera_from_view(vd)
    r = Renderer(0)
    try:
        r.upload_scene(sc)
        eye = np.asarray(vd.eye, np.float64)
        fwd = np.asarray(vd.target, np.float64) - eye
        fwd /= np.linalg.norm(fwd)
        front = np.eye(4, dtype=np.float32)
        front[:3, 3] = eye + 3.0 * fwd
        behind = np.eye(4, dtype=np.float32)
        behind[:3, 3] = eye - 50.0 * fwd
        xf = np.stack([behind.T.ravel(), front.T.ravel()]).astype(np.float32)   # column-major
        r.set_transforms(xf)
        one = np.zeros(1, DRAW_DTYPE)
        one["transform_index"] = 1
        r.set_draws(one)
        view = r.create_view(w, h)
        flags = N.CULL_FRUSTUM | N.RESOLVE_BARY | N.NO_BACKFACE
        view.render(cam, flags)
        ref_keys = view.keys()
        ref_ids = oracle.keys_ids(ref_keys)
        assert (ref_ids != 0xFFFFFFFF).sum() > 20, "the test meshlet must cover pixels"
        n = N.MAX_DRAWS
        draws = np.zeros(n, DRAW_DTYPE)
        draws["transform_index"][-1] = 1
        r.set_draws(draws)
        view.render(cam, flags)
        keys = view.keys()
        ids = oracle.keys_ids(keys)
        cov = ids != 0xFFFFFFFF
        assert np.array_equal(cov, ref_ids != 0xFFFFFFFF)
        assert np.all(ids[cov] >> 7 == n - 1)
        assert np.array_equal(ids[cov] & 0x7F, ref_ids[cov] & 0x7F)
        assert np.array_equal(oracle.keys_depth(keys), oracle.keys_depth(ref_keys))
        assert np.array_equal(view.visible(0), np.array([n - 1], np.uint32))
        print(f"{n} draws: draw {n - 1} owns {cov.sum()} px", flush=True)
        view.close()
        too_many = np.zeros(n + 1, DRAW_DTYPE)
        with pytest.raises(VkgvError) as e:
            r.set_draws(too_many)
        assert e.value.code == N.ERROR_LIMIT
    finally:
        r.close()


def test_zero_draw_frames():
    sc = Scene(1, 1, 8).arrays()
    r = Renderer(0)
    try:
        r.load(sc)
        vd = sc.views[0]
        w, h = 200, 120
        vd.width, vd.height = w, h
        cam = camera_from_view(vd)
        views = [r.create_view(w, h), r.create_view(w // 2, h)]
        views[0].render(cam, FLAGS)
        assert (views[0].color() != 0).any()          # something drawn before
        r.set_draws(np.zeros(0, DRAW_DTYPE))
        for flags in (FLAGS, N.MODE_REFERENCE):
            views[0].render(cam, flags)
            assert np.all(views[0].keys() == np.uint64(N.KEY_CLEAR))
            assert np.all(views[0].color() == 0)
            st = views[0].stats()
            assert st.visible_early == 0 and st.visible_late == 0 and st.draws == 0
            assert len(views[0].visible(0)) == 0
        b = r.create_batch(views)
        b.render([cam, cam], FLAGS | N.GRAPH)
        for v in views:
            assert np.all(v.keys() == np.uint64(N.KEY_CLEAR))
            assert np.all(v.color() == 0)
        b.close()
    finally:
        r.close()
