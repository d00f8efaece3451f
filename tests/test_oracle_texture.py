"""The oracle's textured resolve (oracle/cpuref.c), pinned by closed forms on
a screen-filling quad whose interpolated uvs put pixel centres on texel
centres:

- identity transform: the image is the texture itself (sRGB decode, then
  encode, returns the byte), which fixes the uv interpolation, the texel
  addressing (row 0 at v = 0) and the alpha channel;
- the wrap modes under KHR_texture_transform offsets and a negative scale:
  REPEAT flips, MIRRORED_REPEAT restores, CLAMP_TO_EDGE repeats the edge;
- a quarter-turn uv rotation transposes the texture;
- bilinear weights: a 2-texel black/white texture gives the linear ramp
  clamp(2u - 1/2, 0, 1) before encoding;
- albedo factor and vertex colour multiply the sample.
"""
import numpy as np
import pytest

import oracle
from texscene import quad_scene
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import Texture

W, H = 64, 48
FLAGS = N.CULL_FRUSTUM | N.RESOLVE_TEXTURED


def render(sc, cam, flags=FLAGS):
    rs = oracle.Scene(sc)
    _, _, total = oracle.pyramid_info(W, H)
    keys, rgba, _ = oracle.frame_reference(rs, cam, W, H, flags, np.zeros(total, np.float32))
    assert (keys != np.uint64(0xFFFFFFFFFFFFFFFF)).all()          # the quad covers every pixel
    return rgba.view(np.uint8).reshape(H, W, 4)


def close(got, want, exact_frac=0.999):
    d = np.abs(got.astype(np.int16) - want.astype(np.int16))
    assert d.max() <= 1, d.max()
    assert (d == 0).mean() >= exact_frac, (d == 0).mean()


@pytest.fixture(scope="module")
def tex():
    return np.random.default_rng(1).integers(0, 256, (H, W, 4)).astype(np.uint8)


def test_identity_reproduces_texture(tex):
    sc, cam = quad_scene(W, H, Texture(tex))
    close(render(sc, cam), tex)


def test_flat_and_bary_ignore_texture(tex):
    sc, cam = quad_scene(W, H, Texture(tex))
    img = render(sc, cam, N.CULL_FRUSTUM | N.RESOLVE_BARY)
    assert (img == 255).all()


@pytest.mark.parametrize("wrap,scale,offset,expect", [
    (N.WRAP_REPEAT, (1.0, 1.0), (1.0, -2.0), lambda t: t),                       # whole periods
    (N.WRAP_REPEAT, (-1.0, 1.0), (2.0, 0.0), lambda t: t[:, ::-1]),              # flip
    (N.WRAP_MIRRORED_REPEAT, (-1.0, 1.0), (2.0, 0.0), lambda t: t),              # flip of the mirrored period
    (N.WRAP_MIRRORED_REPEAT, (1.0, 1.0), (1.0, 0.0), lambda t: t[:, ::-1]),      # second period is mirrored
    (N.WRAP_CLAMP_TO_EDGE, (1.0, 1.0), (0.5, 0.0),
     lambda t: t[:, np.minimum(np.arange(W) + W // 2, W - 1)]),                  # edge column repeats
])
def test_wrap_modes(tex, wrap, scale, offset, expect):
    sc, cam = quad_scene(W, H, Texture(tex, wrap, wrap), uv_offset=offset, uv_scale=scale)
    close(render(sc, cam), expect(tex))


def test_quarter_turn_transposes():
    """rotation pi/2, offset (0, 1): tu = v, tv = 1 - u, so pixel (px, py)
    reads texel (x = py, y = S - 1 - px) of an S x S texture."""
    S = 48
    global W, H
    w0, h0 = W, H
    W = H = S
    try:
        t = np.random.default_rng(2).integers(0, 256, (S, S, 4)).astype(np.uint8)
        sc, cam = quad_scene(S, S, Texture(t), uv_offset=(0.0, 1.0), uv_rotation=np.pi / 2)
        img = render(sc, cam)
        px, py = np.meshgrid(np.arange(S), np.arange(S))
        close(img, t[S - 1 - px, py], 0.99)
    finally:
        W, H = w0, h0


def srgb_encode(v):
    v = np.asarray(v, np.float64)
    enc = np.where(v < 0.0031308, v * 12.92, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055)
    return np.floor(np.clip(enc, 0, 1) * 255 + 0.5).astype(np.int16)      # unorm8: round half up


def test_bilinear_ramp():
    t = np.array([[[0, 0, 0, 255], [255, 255, 255, 255]]], np.uint8)          # 1 x 2: black, white
    sc, cam = quad_scene(W, H, Texture(t, N.WRAP_CLAMP_TO_EDGE, N.WRAP_CLAMP_TO_EDGE))
    img = render(sc, cam)
    u = (np.arange(W) + 0.5) / W
    lin = np.clip(2 * u - 0.5, 0, 1)                                             # weight of the white texel
    want = srgb_encode(lin)
    for c in range(3):
        d = np.abs(img[:, :, c].astype(np.int16) - want[None, :])
        assert d.max() <= 1
    assert (img[:, :, 3] == 255).all()
    assert (img[:, :W // 4 - 1, :3] == 0).all() and (img[:, -(W // 4 - 1):, :3] == 255).all()


def test_albedo_and_vertex_colour_multiply(tex):
    albedo = (0.55, 0.3, 1.0, 0.6)              # no exact half-LSB ties
    sc, cam = quad_scene(W, H, Texture(tex), albedo=albedo)
    img = render(sc, cam)
    dec = lambda b: np.where(b / 255 < 0.04045, b / 255 / 12.92, ((b / 255 + 0.055) / 1.055) ** 2.4)
    want = np.empty_like(tex, dtype=np.int16)
    for c in range(3):
        want[..., c] = srgb_encode(dec(tex[..., c].astype(np.float64)) * albedo[c])
    want[..., 3] = np.floor(tex[..., 3] / 255 * albedo[3] * 255 + 0.5)
    close(img, want, 0.99)
