"""vkgv_view_aa_image_noalloc against vkgv_view_aa_image through the C ABI:
the same antialiased image, bit for bit, on two calls in a row (the second on
a smaller view, so the fixed scratch keeps the larger call's contents beyond
the smaller view's pixels), and views above the scratch's 3840 x 2160 pixels
refused with VKGV_ERROR_LIMIT."""
import numpy as np
import pytest

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import Scene, camera_from_view

pytestmark = pytest.mark.gpu


def _image_pair(r, sc, width, height):
    import torch
    vd = sc.views[0]
    vd.width, vd.height = width, height
    cam = camera_from_view(vd)
    view = r.create_view(width, height)
    try:
        flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY
        for _ in range(2):
            view.render(cam, flags)
        ref = view.aa_image(cam)
        got = torch.full_like(ref, float("nan"))
        view.aa_image_noalloc(cam, got)
        torch.cuda.synchronize()
        return ref.cpu().numpy(), got.cpu().numpy()
    finally:
        view.close()


def test_noalloc_image_equals_allocating_entry_point():
    sc = Scene(1, 1, 64).arrays()
    r = Renderer(0)
    try:
        r.load(sc)
        for (w, h) in ((1280, 720), (320, 180)):
            print(f"aa_image_noalloc vs aa_image at {w}x{h}", flush=True)
            ref, got = _image_pair(r, sc, w, h)
            assert ref.shape == (h, w, 3)
            assert np.count_nonzero(ref) > 0, "empty image: the comparison would prove nothing"
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (w, h)
    finally:
        r.close()


def test_noalloc_refuses_views_above_the_scratch():
    import torch
    sc = Scene(1, 1, 8).arrays()
    r = Renderer(0)
    try:
        r.load(sc)
        w, h = 3840, 2176                       # 16 rows more than 4K UHD
        vd = sc.views[0]
        vd.width, vd.height = w, h
        cam = camera_from_view(vd)
        view = r.create_view(w, h)
        view.render(cam, N.CULL_FRUSTUM | N.RESOLVE_BARY)
        out = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
        with pytest.raises(VkgvError) as e:
            view.aa_image_noalloc(cam, out)
        assert e.value.code == N.ERROR_LIMIT
        view.close()
    finally:
        r.close()
