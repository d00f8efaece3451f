"""A scene loaded from glTF (quantized positions, node hierarchy, materials)
rendered by the HIP frame, against the oracle on the same arrays: visible
sets, keys, colour and pyramid, reference and two-pass culling."""
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


@pytest.fixture(scope="module")
def loaded(tmp_path_factory):
    p = tmp_path_factory.mktemp("gltf") / "spheres.glb"
    instanced_spheres(grid=12, spacing=2.2, seed=11).write_glb(p)
    sc = gltf.load_gltf(str(p))
    r = Renderer()
    r.load(sc)
    yield sc, r
    r.close()


def test_loaded_scene_reference_frames(loaded):
    sc, r = loaded
    W, H = 640, 360
    vd = view_desc(W, H, (0.0, 6.0, 18.0), (0.0, -1.0, 0.0), fovy_deg=60.0)
    cam = camera_from_view(vd)
    v = r.create_view(W, H)
    try:
        rs = oracle.Scene(sc)
        _, _, total = oracle.pyramid_info(W, H)
        pyr = np.zeros(total, np.float32)
        for frame in range(2):
            print(f"glTF reference frame {frame}", flush=True)
            rk, rc, rv = oracle.frame_reference(rs, cam, W, H, N.MODE_REFERENCE, pyr)
            v.render(cam, N.MODE_REFERENCE)
            assert np.array_equal(v.visible(0), rv), f"frame {frame}: visible set"
            compare_keys(v.keys(), rk, f"glTF frame {frame}")
            compare_color(v.color(), rc, f"glTF frame {frame}")
            assert np.array_equal(v.pyramid().view(np.uint32), pyr.view(np.uint32))
            assert (oracle.keys_ids(rk) != 0xFFFFFFFF).mean() > 0.2
    finally:
        v.close()


def test_loaded_scene_two_pass(loaded):
    sc, r = loaded
    W, H = 512, 512
    vd = view_desc(W, H, (3.0, 2.0, 16.0), (0.0, -1.0, 0.0), fovy_deg=50.0)
    cam = camera_from_view(vd)
    v = r.create_view(W, H)
    flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS
    try:
        rs = oracle.Scene(sc)
        hist = np.zeros(len(sc.draws), np.uint8)
        for frame in range(2):
            print(f"glTF two-pass frame {frame}", flush=True)
            rk, rc, rp, re, rl = oracle.frame_two_pass(rs, cam, W, H, flags, hist)
            v.render(cam, flags)
            assert np.array_equal(v.visible(0), re)
            assert np.array_equal(v.visible(1), rl)
            compare_keys(v.keys(), rk, f"glTF two-pass frame {frame}")
            compare_color(v.color(), rc, f"glTF two-pass frame {frame}")
    finally:
        v.close()
