"""The HIP path against the golden fixtures (tests/golden/), which are computed
from the definitions, not from the oracle: Hi-Z through the MIN sampler's
tent weights, exact rational triangle coverage, IEC 61966-2-1 sRGB codes.
Bit-exact for the integer / min results; sRGB codes exact away from a
rounding tie and within 1 LSB there."""
import os

import numpy as np
import pytest

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer

from helpers import tiny_scene

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CLEAR = np.uint64(0xFFFFFFFFFFFFFFFF)


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


def identity_camera():
    cam = N.Camera()
    eye = np.eye(4, dtype=np.float32).ravel()
    for f in ("view_projection", "occlusion_view_projection", "prev_view_projection", "prev_occlusion_view_projection"):
        getattr(cam, f)[:] = list(eye)
    return cam


def test_hiz_matches_sampler_definition(rend):
    g = np.load(os.path.join(GOLDEN, "hiz.npz"))
    tri = tiny_scene([(np.zeros((3, 3), np.float32), [0, 1, 2])], [(0, np.eye(4), (1, 1, 1, 1), True)])
    rend.load(tri)
    for k in g.files:
        if not k.startswith("depth_"):
            continue
        d = g[k]
        h, w = d.shape
        print(f"hiz golden {w}x{h}", flush=True)
        v = rend.create_view(w, h)
        try:
            v.write_depth(d)
            v.build_pyramid()
            got = v.pyramid()
            assert np.array_equal(got.view(np.uint32), g["pyramid_" + k[len("depth_"):]].view(np.uint32)), k
        finally:
            v.close()


def test_coverage_matches_exact_rational_golden(rend):
    """Each fixture triangle, placed with an identity view-projection (clip =
    NDC, w = 1), covers exactly the pixel centres of its rational mask."""
    g = np.load(os.path.join(GOLDEN, "raster.npz"))
    W, H = int(g["width"]), int(g["height"])
    meshes, inst = [], []
    for k, t in enumerate(g["tris"]):
        pos = np.array([[x / (W / 2) - 1.0, y / (H / 2) - 1.0, 0.5] for (x, y) in t], np.float32)
        meshes.append((pos, [0, 1, 2]))
        inst.append((k, np.eye(4), (1, 1, 1, 1), True))
    sc = tiny_scene(meshes, inst)
    rend.load(sc)
    v = rend.create_view(W, H)
    cam = identity_camera()
    try:
        for k, mask in enumerate(g["masks"]):
            rend.set_draws(sc.draws[k:k + 1])
            v.render(cam, flags=N.NO_BACKFACE | N.SKIP_RESOLVE)
            got = (v.keys() != CLEAR).astype(np.uint8)
            assert np.array_equal(got, mask), k
    finally:
        v.close()


def test_srgb_codes_match_iec_golden(rend):
    """Flat resolve of 4096 one-pixel quads, each with its own albedo from the
    fixture: the RGBA8 code of every channel equals the IEC encode."""
    g = np.load(os.path.join(GOLDEN, "srgb.npz"))
    c, enc, code = g["c"][:4096], g["encoded"][:4096], g["code"][:4096]
    W = H = 64
    quad = np.array([[-0.5, -0.5, 0.5], [0.5, -0.5, 0.5], [0.5, 0.5, 0.5], [-0.5, 0.5, 0.5]], np.float32)
    inst = []
    for i in range(W * H):
        x, y = i % W, i // W
        M = np.diag([2.0 / W, 2.0 / H, 1.0, 1.0])
        M[0, 3] = (x + 0.5) * 2.0 / W - 1.0
        M[1, 3] = (y + 0.5) * 2.0 / H - 1.0
        a = float(c[i])
        inst.append((0, M, (a, a, a, a), True))
    sc = tiny_scene([(quad, [0, 1, 2, 0, 2, 3])], inst)
    rend.load(sc)
    v = rend.create_view(W, H)
    try:
        v.render(identity_camera(), flags=N.NO_BACKFACE)
        rgba = v.color().view(np.uint8).reshape(H * W, 4).astype(np.int64)
    finally:
        v.close()
    frac = enc * 255.0 + 0.5
    clear_of_tie = np.abs(frac - np.round(frac)) > 1e-3
    for ch in range(3):
        assert np.array_equal(rgba[clear_of_tie, ch], code[clear_of_tie].astype(np.int64)), ch
        assert np.abs(rgba[:, ch] - code.astype(np.int64)).max() <= 1
    want_a = np.floor(np.clip(c.astype(np.float64), 0, 1) * 255 + 0.5)
    assert np.abs(rgba[:, 3] - want_a).max() <= 1
