"""Differentiable frame on the GPU: the antialiased linear image of a frame
and its gradient with respect to the instance transforms, computed from the
frame's own visibility buffer, against the FP64 reference (oracle/
framegrad.c, itself pinned by central differences) fed the same keys; the
torch wrapper's gradients equal the direct calls, and an optimisation of
instance translations towards a target lowers the loss."""
import numpy as np
import pytest
import torch

import oracle
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.frame_grad import DifferentiableFrame
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import Scene, camera_from_view, view_desc

pytestmark = pytest.mark.gpu
TWO_PASS = N.CULL_FRUSTUM | N.CULL_TWO_PASS


@pytest.fixture(scope="module")
def rend():
    r = Renderer()
    yield r
    r.close()


def check_view(rend, rs, view, cam, label, seed=0):
    keys = view.keys()
    H, W = keys.shape
    vp = np.asarray(cam.view_projection, np.float32)
    img = view.aa_image(cam)
    ref = oracle.frame_aa_image(rs, vp, keys, W, H)
    err = np.abs(img.cpu().numpy() - ref).max()
    rng = np.random.default_rng(seed)
    g_img = torch.from_numpy(rng.normal(0.0, 1.0, (H, W, 3)).astype(np.float32)).cuda()
    T = rs.sc.transforms.shape[0]
    g = torch.zeros((T, 16), dtype=torch.float64, device="cuda")
    view.aa_backward(cam, g_img, g)
    torch.cuda.synchronize()
    gref = oracle.frame_aa_backward(rs, vp, keys, W, H, g_img.cpu().numpy().astype(np.float64))
    gg = g.cpu().numpy()
    norm = np.linalg.norm(gg - gref) / max(np.linalg.norm(gref), 1e-300)
    ent = (np.abs(gg - gref) / np.maximum(np.abs(gref), 1e-9 * np.abs(gref).max())).max()
    live = int((np.abs(gref).max(1) > 0).sum())
    print(f"{label} {W}x{H}: image max |gpu - fp64| {err:.2e}; transform gradient |g - ref| / |ref| {norm:.2e}, "
          f"max entry rel {ent:.2e}, {live} transforms with a gradient", flush=True)
    assert err < 1e-6
    assert norm < 1e-10 and ent < 1e-8
    assert live > 0


def test_c1_view_matches_reference(rend):
    sc = Scene(1, 1).arrays()
    rend.load(sc)
    rs = oracle.Scene(sc)
    v0 = sc.views[0]
    view = rend.create_view(v0.width, v0.height)
    try:
        cam = camera_from_view(v0)
        view.render(cam, TWO_PASS)
        check_view(rend, rs, view, cam, "config 1")
    finally:
        view.close()


def test_uneven_views_match_reference(rend):
    sc = Scene(5, 5, 4).arrays()
    rend.load(sc)
    rs = oracle.Scene(sc)
    order = sorted(range(len(sc.views)), key=lambda i: sc.views[i].width * sc.views[i].height)[:2]
    vds = [sc.views[i] for i in order]
    views = [rend.create_view(v.width, v.height) for v in vds]
    batch = rend.create_batch(views)
    try:
        cams = [camera_from_view(v) for v in vds]
        batch.render(cams, TWO_PASS)
        for i, (v, c) in enumerate(zip(views, cams)):
            check_view(rend, rs, v, c, f"config 5.4 view {order[i]}", seed=i)
    finally:
        batch.close()
        for v in views:
            v.close()


def test_errors(rend):
    sc = Scene(1, 1).arrays()
    rend.load(sc)
    v0 = sc.views[0]
    view = rend.create_view(64, 64)
    try:
        cam = camera_from_view(v0)
        view.render(cam, TWO_PASS)
        g_img = torch.zeros((64, 64, 3), dtype=torch.float32, device="cuda")
        small = torch.zeros((3, 16), dtype=torch.float64, device="cuda")
        with pytest.raises(VkgvError):          # fewer transforms than the scene holds
            view.aa_backward(cam, g_img, small)
        with pytest.raises(ValueError):
            view.aa_image(cam, torch.zeros((64, 64, 4), dtype=torch.float32, device="cuda"))
    finally:
        view.close()


def test_autograd_wrapper_and_optimisation(rend):
    """Instance translations perturbed by up to 0.3 m are pulled back towards
    the target frame's by the transform gradient."""
    sc = Scene(1, 1).arrays()
    rend.load(sc)
    v0 = sc.views[0]
    vds = [view_desc(320, 180, tuple(v0.eye), tuple(v0.target), fovy_deg=float(np.degrees(v0.fovy))),
           view_desc(256, 256, (40.0, 30.0, 60.0), (0.0, 0.0, 0.0), fovy_deg=50.0)]
    views = [rend.create_view(v.width, v.height) for v in vds]
    cams = [camera_from_view(v) for v in vds]
    frame = DifferentiableFrame(rend, views, cams)
    try:
        true = torch.from_numpy(sc.transforms.astype(np.float32)).cuda()
        with torch.no_grad():
            targets = [t.clone() for t in frame(true)]
        rng = np.random.default_rng(7)
        start = sc.transforms.astype(np.float32).copy()
        start[:, 12:15] += rng.uniform(-0.3, 0.3, (start.shape[0], 3)).astype(np.float32)
        x = torch.from_numpy(start).cuda().requires_grad_()
        # the wrapper's gradient equals the direct calls
        imgs = frame(x)
        loss = sum(((i - t) ** 2).sum() for i, t in zip(imgs, targets))
        loss.backward()
        direct = torch.zeros((x.shape[0], 16), dtype=torch.float64, device="cuda")
        for v, c, i, t in zip(views, cams, imgs, targets):
            v.aa_backward(c, (2.0 * (i - t)).detach().contiguous(), direct)
        np.testing.assert_allclose(x.grad.cpu().numpy(), direct.float().cpu().numpy(), rtol=1e-6, atol=1e-6)
        # translations only, so the rasterization changes little between steps
        mask = torch.zeros_like(x)
        mask[:, 12:15] = 1.0
        opt = torch.optim.Adam([x], lr=0.01)
        losses = []
        for _ in range(30):
            opt.zero_grad()
            imgs = frame(x)
            loss = sum(((i - t) ** 2).sum() for i, t in zip(imgs, targets))
            loss.backward()
            x.grad *= mask
            opt.step()
            losses.append(loss.item())
        err0 = np.abs(start[:, 12:15] - sc.transforms[:, 12:15]).mean()
        err1 = np.abs(x.detach().cpu().numpy()[:, 12:15] - sc.transforms[:, 12:15]).mean()
        print(f"optimisation: loss {losses[0]:.3f} -> {min(losses):.3f}; mean translation error {err0:.4f} -> {err1:.4f}")
        assert min(losses) < 0.7 * losses[0]
    finally:
        frame.close()
        for v in views:
            v.close()
