"""Every raster tuning gives the same frame: keys, colours and visible lists
of single-view frames and of a captured-graph batch frame are bit-identical
to those of the default tuning (which the parity tests check against the
oracle), also when a low small-triangle extent fills the big-triangle queues
and their triangles take the direct path."""
import numpy as np
import pytest

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import Scene, camera_from_view

pytestmark = pytest.mark.gpu

DEFAULT = (16, 3, 8, 8192)
TUNINGS = [(4, 2, 16, 0), (8, 3, 32, 8192), (32, 1, 1, 0), (16, 16, 256, 100), (64, 3, 16, 1)]
FLAGS = [N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY, N.MODE_REFERENCE]


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


def frames(view, cam, flags):
    """Two frames from a cleared history and an all-far pyramid."""
    out = []
    view.reset_history()
    view.write_pyramid(np.zeros(view.pyramid_texels, np.float32))
    for _ in range(2):
        view.render(cam, flags)
        out.append((view.keys(), view.color(), np.sort(view.visible(0)), np.sort(view.visible(1))))
    return out


@pytest.mark.parametrize("config", [1, 2])
def test_tunings_give_identical_frames(rend, config):
    sc = Scene(config, config).arrays()
    rend.load(sc)
    v0 = sc.views[0]
    view = rend.create_view(v0.width, v0.height)
    cam = camera_from_view(v0)
    try:
        for flags in FLAGS:
            rend.set_tuning(*DEFAULT)
            ref = frames(view, cam, flags)
            for t in TUNINGS:
                rend.set_tuning(*t)
                got = frames(view, cam, flags)
                assert view.stats().overflow == 0, t
                for f, (a, b) in enumerate(zip(ref, got)):
                    for k in range(4):
                        assert np.array_equal(a[k], b[k]), (config, flags, t, f, k)
            print(f"config {config} flags {flags:#x}: {len(TUNINGS)} tunings identical to the default", flush=True)
    finally:
        rend.set_tuning(*DEFAULT)
        view.close()


def test_batch_graph_recaptures_on_tuning_change(rend):
    sc = Scene(3, 3).arrays()
    rend.load(sc)
    vds = [sc.views[i] for i in (0, 5, 20, 41)]
    views = [rend.create_view(v.width, v.height) for v in vds]
    batch = rend.create_batch(views)
    cams = [camera_from_view(v) for v in vds]
    flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY | N.GRAPH
    try:
        rend.set_tuning(*DEFAULT)
        batch.render(cams, flags)
        ref = [(v.keys(), v.color()) for v in views]
        full = 0
        for t in TUNINGS:
            rend.set_tuning(*t)
            batch.render(cams, flags)
            if batch.stats().overflow:
                # a small-triangle extent below the default sends more triangles to the
                # binned path than its queues hold (2048x2048 cascades): the direct path
                assert t[0] < DEFAULT[0], t
                full += 1
            for i, v in enumerate(views):
                assert np.array_equal(v.keys(), ref[i][0]) and np.array_equal(v.color(), ref[i][1]), (t, i)
        print(f"batch graph: identical frames under {len(TUNINGS)} tunings ({full} with a full queue)", flush=True)
    finally:
        rend.set_tuning(*DEFAULT)
        batch.close()
        for v in views:
            v.close()


def test_tuning_limits(rend):
    with pytest.raises(VkgvError):
        rend.set_tuning(16, 3, 16, 8193)
    with pytest.raises(VkgvError):
        rend.set_tuning(0, 3, 16, 0)
    rend.set_tuning(*DEFAULT)
    assert rend.get_tuning() == DEFAULT


def test_native_autotuner_picks_a_valid_tuning(rend):
    """vkgv_autotune_batch (the tuner C and C++ hosts call): it applies a
    tuning from the candidate set whose frames fill no queue, no slower than
    the default tuning timed the same way (within timing noise), and the frame
    under it equals the default frame bit for bit."""
    sc = Scene(3, 3).arrays()
    rend.load(sc)
    vds = [sc.views[i] for i in (0, 5, 9, 20, 41, 63)]
    views = [rend.create_view(v.width, v.height) for v in vds]
    batch = rend.create_batch(views)
    cams = [camera_from_view(v) for v in vds]
    flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY | N.GRAPH
    try:
        rend.set_tuning(*DEFAULT)
        batch.render(cams, flags)
        ref = [(v.keys(), v.color()) for v in views]
        best, ms = batch.autotune(cams, flags, frames_per_candidate=3)
        print(f"native autotuner: {best} at {ms:.4f} ms per batch frame", flush=True)
        assert rend.get_tuning() == best
        assert best[0] in (4, 8, 16, 32) and best[1] in (2, 3) and best[2] in (8, 16, 32)
        assert 0.0 < ms < 1e29
        batch.render(cams, flags)
        assert batch.stats().overflow == 0
        for i, v in enumerate(views):
            assert np.array_equal(v.keys(), ref[i][0]) and np.array_equal(v.color(), ref[i][1]), i
    finally:
        rend.set_tuning(*DEFAULT)
        batch.close()
        for v in views:
            v.close()
