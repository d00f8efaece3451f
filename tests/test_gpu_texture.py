"""Textured resolve (VKGV_RESOLVE_TEXTURED) on the HIP path against the oracle
on the same inputs: the screen-filling quad under each wrap mode and a uv
rotation, a textured glTF scene (REPEAT / MIRRORED_REPEAT / CLAMP_TO_EDGE
samplers, KHR_texture_transform, an untextured material) in single-view and
batched frames, and the state error when a material names a texture that
was not uploaded."""
import numpy as np
import pytest

import oracle
from test_gpu_parity import compare_color, compare_keys
from texscene import quad_scene, textured_spheres
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd import gltf
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import Texture, camera_from_view, view_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


def exact(gpu_rgba, ref_rgba, label):
    same = np.array_equal(gpu_rgba, ref_rgba)
    print(f"{label}: colour exact={same}", flush=True)
    compare_color(gpu_rgba, ref_rgba, label)
    assert same, label


@pytest.mark.parametrize("case", ["identity", "repeat", "mirror", "clamp", "rotate"])
def test_quad_matches_oracle(rend, case):
    W, H = 320, 200
    rng = np.random.default_rng(7)
    tex = rng.integers(0, 256, (H // 2, W // 3, 4)).astype(np.uint8)
    wrap = {"identity": 0, "repeat": N.WRAP_REPEAT, "mirror": N.WRAP_MIRRORED_REPEAT,
            "clamp": N.WRAP_CLAMP_TO_EDGE, "rotate": N.WRAP_REPEAT}[case]
    kw = {"identity": {}, "repeat": dict(uv_offset=(0.3, -1.7), uv_scale=(2.5, 1.5)),
          "mirror": dict(uv_offset=(-0.4, 0.2), uv_scale=(3.0, -2.0)),
          "clamp": dict(uv_offset=(0.35, -0.25), uv_scale=(1.5, 1.5)),
          "rotate": dict(uv_offset=(0.1, 0.9), uv_scale=(1.3, 0.8), uv_rotation=0.7)}[case]
    sc, cam = quad_scene(W, H, Texture(tex, wrap, wrap), **kw)
    print(f"quad {case}", flush=True)
    rend.load(sc)
    v = rend.create_view(W, H)
    try:
        flags = N.CULL_FRUSTUM | N.RESOLVE_TEXTURED
        rs = oracle.Scene(sc)
        _, _, total = oracle.pyramid_info(W, H)
        rk, rc, _ = oracle.frame_reference(rs, cam, W, H, flags, np.zeros(total, np.float32))
        v.render(cam, flags)
        compare_keys(v.keys(), rk, f"quad {case}")
        exact(v.color(), rc, f"quad {case}")
    finally:
        v.close()


@pytest.fixture(scope="module")
def spheres(tmp_path_factory):
    p = tmp_path_factory.mktemp("tex") / "spheres.glb"
    textured_spheres(grid=8, seed=9).write_glb(p)
    return gltf.load_gltf(str(p))


def test_textured_gltf_frames(rend, spheres):
    sc = spheres
    rend.load(sc)
    rs = oracle.Scene(sc)
    for k, (W, H, eye) in enumerate([(640, 360, (0.0, 7.0, 14.0)), (333, 257, (4.0, 2.5, 9.0))]):
        vd = view_desc(W, H, eye, (0.0, 0.0, 0.0), fovy_deg=55.0)
        cam = camera_from_view(vd)
        v = rend.create_view(W, H)
        try:
            for flags in (N.CULL_FRUSTUM | N.RESOLVE_TEXTURED,
                          N.CULL_FRUSTUM | N.RESOLVE_TEXTURED | N.RESOLVE_BARY | N.CULL_CONE):
                print(f"textured glTF view {k} flags {flags:#x}", flush=True)
                _, _, total = oracle.pyramid_info(W, H)
                rk, rc, rv = oracle.frame_reference(rs, cam, W, H, flags, np.zeros(total, np.float32))
                v.render(cam, flags)
                assert np.array_equal(v.visible(0), rv)
                compare_keys(v.keys(), rk, f"textured view {k}")
                exact(v.color(), rc, f"textured view {k}")
                covered = oracle.keys_ids(rk) != 0xFFFFFFFF
                assert covered.mean() > 0.2
                # the textures make the image vary within materials
                assert len(np.unique(rc[covered])) > 1000
        finally:
            v.close()


def test_textured_batch_matches_single_views(rend, spheres):
    sc = spheres
    rend.load(sc)
    specs = [(256, 192, (0.0, 6.0, 12.0)), (200, 300, (-5.0, 3.0, 8.0)), (480, 270, (6.0, 4.0, -9.0))]
    views = [rend.create_view(W, H) for W, H, _ in specs]
    cams = [camera_from_view(view_desc(W, H, e, (0.0, 0.0, 0.0), fovy_deg=50.0)) for W, H, e in specs]
    batch = rend.create_batch(views)
    flags = N.CULL_FRUSTUM | N.RESOLVE_TEXTURED
    try:
        print("textured batch", flush=True)
        batch.render(cams, flags)
        rend.sync()
        got = [v.color().copy() for v in views]
        for v, c, g in zip(views, cams, got):
            v.render(c, flags)
            assert np.array_equal(v.color(), g)
    finally:
        batch.close()
        for v in views:
            v.close()


def test_missing_textures_is_a_state_error(rend, spheres):
    sc = spheres
    rend.load(sc)
    rend.upload_textures(sc.textures[:1])          # materials name textures 0..2
    v = rend.create_view(64, 64)
    try:
        cam = camera_from_view(view_desc(64, 64, (0.0, 5.0, 10.0), (0.0, 0.0, 0.0)))
        with pytest.raises(VkgvError, match="texture"):
            v.render(cam, N.CULL_FRUSTUM | N.RESOLVE_TEXTURED)
        v.render(cam, N.CULL_FRUSTUM | N.RESOLVE_BARY)   # untextured frames still run
        rend.upload_textures(sc.textures)
        v.render(cam, N.CULL_FRUSTUM | N.RESOLVE_TEXTURED)
        rend.sync()
    finally:
        v.close()
