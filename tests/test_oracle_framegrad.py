"""The FP64 reference of the differentiable frame (oracle/framegrad.c), pinned
on the CPU:

- on a one-material scene its antialiased image equals the differentiable
  form's FP64 forward (oracle/diffref.c, itself pinned by closed forms) fed
  the same ids, which pins the geometry lookup through the visibility-buffer
  id (draw -> meshlet -> vertices, instance transform, camera);
- its gradient with respect to the instance transforms equals
  Ridders-extrapolated central differences of the loss, with the keys and
  the antialiasing state held fixed, on seeded transform entries of a
  256-instance, 256-material scene.
"""
import numpy as np
import pytest

import oracle
from diffcheck import ridders
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import Scene, camera_from_view, view_desc


def small_view(v, w, h):
    return view_desc(w, h, tuple(v.eye), tuple(v.target), fovy_deg=float(np.degrees(v.fovy)), up=tuple(v.up),
                     znear=v.znear, zfar=v.zfar)


def frame_keys(rs, cam, w, h):
    _, _, total = oracle.pyramid_info(w, h)
    keys, _, _ = oracle.frame_reference(rs, cam, w, h, N.CULL_FRUSTUM, np.zeros(total, np.float32))
    return keys


def test_image_equals_differentiable_form_on_one_material():
    sc = Scene(4, 4).arrays()
    assert len(sc.materials) >= 1 and len(sc.transforms) == 1
    rs = oracle.Scene(sc)
    W, H = 96, 96
    view = small_view(sc.views[0], W, H)
    cam = camera_from_view(view)
    keys = frame_keys(rs, cam, W, H)
    vp = np.asarray(cam.view_projection, np.float32)
    img = oracle.frame_aa_image(rs, vp, keys, W, H)
    # the same frame in the differentiable form's terms: one mesh, the
    # meshlets' triangles concatenated, ids i * T + t + 1
    m = sc.meshes[0]
    tris, base = [], []
    for ml in m.meshlets:
        base.append(sum(len(t) for t in tris))
        vo, to, tc = int(ml["vertex_offset"]), int(ml["triangle_offset"]), int(ml["triangle_count"])
        loc = m.triangle_indices[to:to + 3 * tc].reshape(-1, 3).astype(np.int64)
        tris.append(m.vertex_indices[vo + loc])
    tris = np.concatenate(tris).astype(np.int32)
    base = np.asarray(base)
    covered = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
    fid = (np.uint64(0xFFFFFFFE) - (keys & np.uint64(0xFFFFFFFF))).astype(np.int64)
    draw, tri = fid >> 7, fid & 127
    ids = np.where(covered, base[sc.draws["meshlet_index"][np.where(covered, draw, 0)]] + tri + 1, 0).astype(np.int32)
    depth = np.where(covered, oracle.keys_depth(keys.ravel()).reshape(H, W), 0.0).astype(np.float32)
    M = sc.transforms[0].reshape(4, 4).T.astype(np.float64)          # column-major storage -> row-major
    VP = vp.reshape(4, 4).T.astype(np.float64)
    pos = m.vertices["position"].astype(np.float64)
    clip = np.einsum("ab,bc,vc->va", VP, M, np.c_[pos, np.ones(len(pos))])[None, None]
    albedo = np.asarray(sc.materials["albedo_factor"][int(sc.primitives["material_index"][0])][:3], np.float64)
    attr = np.broadcast_to(albedo, (len(pos), 3)).copy()
    _, ref = oracle.diff_forward(clip, attr, tris, ids[None], depth[None])
    assert covered.mean() > 0.2
    edge = np.abs(img - albedo * covered[..., None]).max(-1) > 1e-9
    assert edge.sum() > 50                      # the silhouette is antialiased
    np.testing.assert_allclose(img, ref[0], atol=1e-12)


@pytest.fixture(scope="module")
def c1_small():
    sc = Scene(1, 1).arrays()
    rs = oracle.Scene(sc)
    W, H = 256, 144
    view = small_view(sc.views[0], W, H)
    cam = camera_from_view(view)
    keys = frame_keys(rs, cam, W, H)
    vp = np.asarray(cam.view_projection, np.float32)
    return sc, rs, W, H, cam, vp, keys


def test_transform_gradient_matches_central_differences(c1_small):
    sc, rs, W, H, cam, vp, keys = c1_small
    img = oracle.frame_aa_image(rs, vp, keys, W, H)
    rng = np.random.default_rng(21)
    target = img + rng.normal(0.0, 0.05, img.shape)
    g = oracle.frame_aa_backward(rs, vp, keys, W, H, 2.0 * (img - target))
    xf = sc.transforms.astype(np.float64).reshape(-1, 16)
    # transforms whose gradient is not zero: instances with silhouette pairs
    live = np.nonzero(np.abs(g).max(1) > 0)[0]
    assert len(live) > 20
    params = [(int(rng.choice(live)), int(c) * 4 + int(r)) for c, r in zip(rng.integers(0, 4, 24), rng.integers(0, 3, 24))]
    pix = np.arange(W * H, dtype=np.int64)
    errs = []
    for t, k in params:
        state = np.zeros((pix.size, 4), np.int8)
        oracle.frame_aa_loss_at(rs, vp, keys, W, H, xf, target, pix, state, 1)

        def f(s):
            x = xf.copy()
            x[t, k] += s
            return oracle.frame_aa_loss_at(rs, vp, keys, W, H, x, target, pix, state, 2)
        d, _ = ridders(f, 1e-3)
        errs.append(abs(d - g[t, k]) / max(abs(d), 1e-6 * np.abs(g).max()))
    assert max(errs) < 1e-6, errs


def test_loss_at_equals_image(c1_small):
    sc, rs, W, H, cam, vp, keys = c1_small
    img = oracle.frame_aa_image(rs, vp, keys, W, H)
    target = np.zeros_like(img)
    pix = np.random.default_rng(2).permutation(W * H).astype(np.int64)
    loss = oracle.frame_aa_loss_at(rs, vp, keys, W, H, None, target, pix)
    assert loss == pytest.approx(float((img ** 2).sum()), rel=1e-12)
