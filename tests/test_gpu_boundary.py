"""Boundary variants that take device memory (the reference's primitives hold
buffer device addresses, mesh_common.h.glsl:102-113):
vkgv_upload_scene_device and vkgv_set_transforms_device give frames equal,
bit for bit, to the host-copy entry points; malformed meshes in device
memory are refused by the device-side check and leave frames refused."""
import numpy as np
import pytest

from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.renderer import Renderer, VkgvError
from vk_gltf_viewer_amd.scene import Scene, camera_from_view

pytestmark = pytest.mark.gpu

FLAGS = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.CULL_CONE | N.RESOLVE_BARY


def _frames(r, sc, n=2):
    vd = sc.views[0]
    cam = camera_from_view(vd)
    eye = tuple(vd.eye)
    v = r.create_view(vd.width, vd.height)
    out = []
    try:
        for _ in range(n):
            v.render(cam, FLAGS, eye=eye)
            out.append((v.keys(), v.color(), v.visible(0), v.visible(1)))
    finally:
        v.close()
    return out


def test_device_upload_equals_host_upload():
    sc = Scene(1, 1, 128).arrays()
    assert any(m.cones is not None for m in sc.meshes)
    host, dev = Renderer(0), Renderer(0)
    try:
        host.load(sc)
        dev.upload_scene_device(sc)
        dev.set_transforms(sc.transforms)
        dev.set_draws(sc.draws)
        a, b = _frames(host, sc), _frames(dev, sc)
        for f, (x, y) in enumerate(zip(a, b)):
            for k in range(4):
                assert np.array_equal(x[k], y[k]), (f, k)
        covered = int((a[-1][0] != np.uint64(N.KEY_CLEAR)).sum())
        print(f"device upload: {len(sc.meshes)} meshes, frames bit-identical ({covered} px covered)", flush=True)
        assert covered > 1000
    finally:
        host.close()
        dev.close()


@pytest.mark.parametrize("what", ["triangle index", "vertex index", "meshlet range"])
def test_device_upload_refuses_malformed_meshes(what):
    sc = Scene(1, 1, 4).arrays()
    r = Renderer(0)
    try:
        r.load(sc)
        vd = sc.views[0]
        cam = camera_from_view(vd)
        v = r.create_view(160, 90)
        v.render(cam, FLAGS)

        def corrupt(i, parts):
            if i != 0:
                return
            if what == "triangle index":
                parts[1][7] = 200                       # a triangle byte beyond any meshlet's 64 vertices
            elif what == "vertex index":
                parts[0][4:8] = 0xFF                    # vertex index 2^32 - 1 of the first meshlet
            else:
                parts[3][0:4] = 0xFF                    # the first meshlet's vertex offset past the index array

        with pytest.raises(VkgvError) as e:
            r.upload_scene_device(sc, corrupt=corrupt)
        assert e.value.code == N.ERROR_INVALID_ARGUMENT and "mesh 0" in str(e.value), str(e.value)
        with pytest.raises(VkgvError) as e2:
            v.render(cam, FLAGS)
        assert e2.value.code == N.ERROR_STATE
        v.close()
    finally:
        r.close()


def test_set_transforms_device_equals_host():
    import torch
    sc = Scene(1, 1, 64).arrays()
    rng = np.random.default_rng(3)
    host, dev = Renderer(0), Renderer(0)
    try:
        for r in (host, dev):
            r.upload_scene(sc)
            r.set_transforms(sc.transforms)
            r.set_draws(sc.draws)
        vd = sc.views[0]
        cam = camera_from_view(vd)
        vh, vdv = host.create_view(vd.width, vd.height), dev.create_view(vd.width, vd.height)
        s = torch.cuda.Stream()
        for f in range(3):
            t = sc.transforms.reshape(-1, 16).copy()
            t[:, 12:15] += rng.normal(0, 0.5, (len(t), 3)).astype(np.float32)   # move every instance
            host.set_transforms(t)
            with torch.cuda.stream(s):
                td = torch.from_numpy(t).to("cuda", non_blocking=False)
                dev.set_transforms_device(td)
                vdv.render(cam, FLAGS | N.MOTION_VECTORS, stream=s.cuda_stream)
            s.synchronize()
            vh.render(cam, FLAGS | N.MOTION_VECTORS)
            assert np.array_equal(vh.keys(), vdv.keys()), f
            assert np.array_equal(vh.color(), vdv.color()), f
            assert np.array_equal(vh.motion(), vdv.motion()), f
        assert np.array_equal(host.read_transforms(len(t)), dev.read_transforms(len(t)))
        vh.close()
        vdv.close()
    finally:
        host.close()
        dev.close()
