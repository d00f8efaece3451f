"""The oracle of the transform update and the motion vectors (oracle/anim.c),
pinned by closed forms:

- the acos / sin series against the math library;
- LINEAR at midpoints, STEP holding the previous key, the fmod loop (a time
  equal to the last key wraps to the first value, as the reference's loop
  does), times before the first key;
- CUBICSPLINE with zero tangents is v0 + (v1 - v0)(3t^2 - 2t^3), with
  tangents the glTF Hermite form in float64;
- slerp between rotations about one axis is the rotation by the interpolated
  angle, on the shortest arc whatever the sign of the second key;
- a later channel on the same node and path wins;
- an unanimated hierarchy gives float64 products of T R S down each chain,
  in depth-first order, and equals the host library's World traversal
  (vkgv_world_build) bit for bit;
- motion vectors: zero for a still frame, dx / (2 w) for an orthographic
  camera moved by dx, -d / (2 w) for an object moved by d.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from animscene import random_hierarchy
from texscene import quad_scene
from vk_gltf_viewer_amd import _native as N
from vk_gltf_viewer_amd.scene import (ANIM_CHANNEL_DTYPE, ANIM_SAMPLER_DTYPE, NODE_DTYPE, Hierarchy, Texture,
                                      camera_advance, view_desc)


def sampler(count, interp, off=0, voff=0):
    return np.array([(off, count, voff, interp)], np.uint32).view(ANIM_SAMPLER_DTYPE)


def test_series_match_libm():
    x = np.linspace(0, 1, 2001)
    got = np.array([oracle.series_acos(v) for v in x])
    assert np.abs(got - np.arccos(x)).max() < 4e-16 * 4
    a = np.linspace(0, np.pi / 2, 2001)
    got = np.array([oracle.series_sin(v) for v in a])
    assert np.abs(got - np.sin(a)).max() < 4e-16


def test_linear_step_and_loop():
    t = np.array([1.0, 3.0], np.float32)
    v = np.array([0, 0, 0, 2, 4, -6], np.float32)
    lin = sampler(2, N.INTERP_LINEAR)
    np.testing.assert_array_equal(oracle.anim_sample(lin, t, v, 3, 2.0), [1, 2, -3])
    np.testing.assert_array_equal(oracle.anim_sample(lin, t, v, 3, 0.5), [0, 0, 0])          # before the first key
    np.testing.assert_array_equal(oracle.anim_sample(lin, t, v, 3, 3.0), [0, 0, 0])          # wraps to time 0
    for k in range(1, 5):
        np.testing.assert_array_equal(oracle.anim_sample(lin, t, v, 3, 2.0 + 3.0 * k), [1, 2, -3])
    ts = np.array([0.5, 1.0, 2.0], np.float32)
    vs = np.arange(9, dtype=np.float32)
    st = sampler(3, N.INTERP_STEP)
    np.testing.assert_array_equal(oracle.anim_sample(st, ts, vs, 3, 1.5), [3, 4, 5])
    np.testing.assert_array_equal(oracle.anim_sample(st, ts, vs, 3, 0.75), [0, 1, 2])
    np.testing.assert_array_equal(oracle.anim_sample(st, ts, vs, 3, 0.2), [0, 1, 2])


def test_cubic_spline():
    t = np.array([0.5, 2.5], np.float32)
    v0, v1 = np.array([1.0, -2.0, 3.0]), np.array([4.0, 0.5, -1.0])
    zero = np.zeros(3)
    vals = np.concatenate([zero, v0, zero, zero, v1, zero]).astype(np.float32)       # (in, value, out) per key
    sp = sampler(2, N.INTERP_CUBICSPLINE)
    for time in (0.6, 1.0, 1.5, 2.2, 2.49):
        u = (time - 0.5) / 2.0
        want = v0 + (v1 - v0) * (3 * u * u - 2 * u ** 3)
        np.testing.assert_allclose(oracle.anim_sample(sp, t, vals, 3, time), want, rtol=0, atol=2e-6)
    np.testing.assert_array_equal(oracle.anim_sample(sp, t, vals, 3, 0.1), v0.astype(np.float32))
    rng = np.random.default_rng(3)
    b0, a1 = rng.normal(size=3), rng.normal(size=3)
    vals = np.concatenate([zero, v0, b0, a1, v1, zero]).astype(np.float32)
    for time in (0.7, 1.3, 2.4):
        u, dt = (time - 0.5) / 2.0, 2.0
        want = ((2 * u ** 3 - 3 * u ** 2 + 1) * v0 + (u ** 3 - 2 * u ** 2 + u) * dt * vals[6:9]
                + (-2 * u ** 3 + 3 * u ** 2) * v1 + (u ** 3 - u ** 2) * dt * vals[9:12])
        np.testing.assert_allclose(oracle.anim_sample(sp, t, vals, 3, time), want, rtol=0, atol=1e-5)


def axis_quat(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.r_[axis * np.sin(ang / 2), np.cos(ang / 2)]


@pytest.mark.parametrize("a0,a1", [(0.0, 1.0), (-0.5, 2.0), (0.3, 0.3001), (2.5, -2.5)])
def test_slerp_is_rotation_by_interpolated_angle(a0, a1):
    axis = [0.3, -0.8, 0.5]
    q0, q1 = axis_quat(axis, a0), axis_quat(axis, a1)
    t = np.array([0.0, 1.0], np.float32)
    sp = sampler(2, N.INTERP_LINEAR)
    for sign in (1, -1):
        vals = np.concatenate([q0, sign * q1]).astype(np.float32)
        for time in (0.1, 0.5, 0.9):
            got = oracle.anim_sample(sp, t, vals, 4, time).astype(np.float64)
            # shortest arc: the angle difference folded into (-pi, pi]
            d = (a1 - a0 + np.pi) % (2 * np.pi) - np.pi
            want = axis_quat(axis, a0 + time * d)
            if np.dot(got, want) < 0:
                want = -want
            np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)


def test_later_channel_wins():
    nodes = np.zeros(1, NODE_DTYPE)
    nodes["mesh"] = 0
    nodes["rotation"] = (0, 0, 0, 1)
    nodes["scale"] = 1
    sp = np.array([(0, 1, 0, 0), (1, 1, 3, 0)], np.uint32).view(ANIM_SAMPLER_DTYPE).ravel()
    times = np.array([1.0, 1.0], np.float32)
    values = np.array([1, 2, 3, 7, 8, 9], np.float32)
    ch = np.array([(0, 0, 0, 0), (0, 1, 0, 0)], np.uint32).view(ANIM_CHANNEL_DTYPE).ravel()
    h = Hierarchy(nodes, np.zeros(0, np.uint32), np.array([0], np.uint32), sp, times, values, ch)
    m = oracle.update_transforms(h, 0.5)[0].reshape(4, 4).T
    np.testing.assert_array_equal(m[:3, 3], [7, 8, 9])


def trs64(n):
    x, y, z, w = np.asarray(n["rotation"], np.float64)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(n["scale"], np.float64)[None, :]
    M[:3, 3] = n["translation"]
    return M


def test_unanimated_hierarchy_closed_form_and_host_world():
    h = random_hierarchy(5, nodes=120, animated_frac=0.0)
    got = oracle.update_transforms(h, 1.25)
    want = []

    def walk(i, M):
        M = M @ trs64(h.nodes[i])
        if h.nodes[i]["mesh"] >= 0:
            want.append(M)
        f, c = int(h.nodes[i]["first_child"]), int(h.nodes[i]["child_count"])
        for k in h.children[f:f + c]:
            walk(int(k), M)
    for r in h.roots:
        walk(int(r), np.eye(4))
    assert got.shape == (len(want), 16)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g.reshape(4, 4).T, w, rtol=1e-5, atol=1e-4)
    # the host library's traversal (vkgv_world_build) gives the same bits
    L = N.host()
    nodes = np.ascontiguousarray(h.nodes)
    ch = np.ascontiguousarray(h.children)
    rt = np.ascontiguousarray(h.roots)
    meshes = np.zeros(8, np.uint32)                      # 4 meshes of 0 primitives
    nd, nt = C.c_uint32(0), C.c_uint32(0)
    args = [nodes.ctypes.data_as(C.POINTER(N.Node)), len(nodes), ch.ctypes.data, rt.ctypes.data, len(rt),
            meshes.ctypes.data, meshes.ctypes.data, meshes.ctypes.data]
    assert L.vkgv_world_build(*args, None, C.byref(nd), None, C.byref(nt)) == 0
    tr = np.zeros((nt.value, 16), np.float32)
    assert L.vkgv_world_build(*args, None, C.byref(nd), tr.ctypes.data, C.byref(nt)) == 0
    assert np.array_equal(tr.view(np.uint32), got.view(np.uint32))


def test_animated_hierarchy_composes_sampled_trs():
    h = random_hierarchy(6, nodes=60, animated_frac=0.7)
    time = 1.7
    trs = {i: (h.nodes[i]["translation"].copy(), h.nodes[i]["rotation"].copy(), h.nodes[i]["scale"].copy())
           for i in range(len(h.nodes))}
    for c in h.channels:
        i, p = int(c["node"]), int(c["path"])
        v = oracle.anim_sample(h.samplers[int(c["sampler"])], h.times, h.values, 4 if p == 1 else 3, time)
        t, q, s = trs[i]
        (t, q, s)[p][:] = v
    got = oracle.update_transforms(h, time)
    k = 0

    def walk(i, M):
        nonlocal k
        n = np.zeros((), NODE_DTYPE)
        n["translation"], n["rotation"], n["scale"] = trs[i]
        M = M @ trs64(n)
        if h.nodes[i]["mesh"] >= 0:
            np.testing.assert_allclose(got[k].reshape(4, 4).T, M, rtol=1e-4, atol=1e-3)
            k += 1
        f, c = int(h.nodes[i]["first_child"]), int(h.nodes[i]["child_count"])
        for j in h.children[f:f + c]:
            walk(int(j), M)
    for r in h.roots:
        walk(int(r), np.eye(4))
    assert k == len(got) > 10


W, H = 64, 48


def motion_case(dx_cam=0.0, d_obj=0.0):
    sc, cam0 = quad_scene(W, H, Texture(np.zeros((1, 1, 4), np.uint8)))
    v1 = view_desc(W, H, (dx_cam, 0.0, 2.0), (dx_cam, 0.0, 0.0), znear=0.5, zfar=4.0, ortho_half=(1.0, 1.0))
    cam = camera_advance(cam0, v1)
    prev = sc.transforms.copy()
    sc.transforms = sc.transforms.copy()
    sc.transforms[0, 12] = d_obj                         # object moved by d along x
    rs = oracle.Scene(sc)
    _, _, total = oracle.pyramid_info(W, H)
    keys, _, _ = oracle.frame_reference(rs, cam, W, H, N.CULL_FRUSTUM, np.zeros(total, np.float32))
    return oracle.motion_vectors(rs, cam, prev, keys, W, H)


def test_motion_vectors_closed_forms():
    assert np.abs(motion_case()).max() < 1e-6
    mv = motion_case(dx_cam=0.25)
    np.testing.assert_allclose(mv[..., 0], 0.25 / 2, rtol=0, atol=1e-6)
    np.testing.assert_allclose(mv[..., 1], 0.0, rtol=0, atol=1e-6)
    mv = motion_case(d_obj=0.2)
    np.testing.assert_allclose(mv[..., 0], -0.2 / 2, rtol=0, atol=1e-6)
    np.testing.assert_allclose(mv[..., 1], 0.0, rtol=0, atol=1e-6)


def test_lower_bound_is_a_bisection():
    """Keys found by bisection (std::lower_bound): on ascending times the
    first key >= time, as a linear scan finds it."""
    rng = np.random.default_rng(12)
    for k in range(1, 40):
        t = np.sort(rng.uniform(0, 5, k)).astype(np.float32)
        v = np.arange(k, dtype=np.float32).repeat(3)
        sp = sampler(k, N.INTERP_STEP)
        for time in rng.uniform(0, float(t[-1]), 20).astype(np.float32):
            i = int(np.searchsorted(t, time, side="left"))
            want = 0 if i == 0 else (k - 1 if i == k else i - 1)
            assert oracle.anim_sample(sp, t, v, 3, time)[0] == want


def _ulp32(x):
    return np.spacing(np.abs(np.float32(x))).astype(np.float64)


def test_cubic_spline_and_slerp_ulp_bounds():
    """The deliberate deviations from the reference's float forms, bounded.
    The reference's cubic spline takes t^2 and t^3 with std::powf
    (assets.hpp:88-89); here (oracle and kernel alike) they are t * t and
    (t * t) * t, the rest of the expression in the reference's order.  Its
    slerp is fastgltf's float acos / sin; here a double-precision series.
    On random keys both stay within a few float32 ulps of the float64 closed
    forms, and the spline within a few ulps of the reference's own float
    expression (powf emulated as the correctly rounded t^3)."""
    rng = np.random.default_rng(17)
    times = np.array([0.0, 1.0], np.float32)
    sp = sampler(2, N.INTERP_CUBICSPLINE)
    worst64 = worst_ref = 0.0
    for _ in range(400):
        vals = rng.normal(0, 2, 6 * 3).astype(np.float32)          # (in, value, out) x 2 keys x 3 comps
        t = np.float32(rng.uniform(0.02, 0.98))
        got = oracle.anim_sample(sp, times, vals, 3, float(t)).astype(np.float64)
        v0, b0, a1, v1 = vals[3:6], vals[6:9], vals[9:12], vals[12:15]
        u = np.float64(t)
        exact = ((2 * u ** 3 - 3 * u ** 2 + 1) * v0 + (u ** 3 - 2 * u ** 2 + u) * b0
                 + (-2 * u ** 3 + 3 * u ** 2) * v1 + (u ** 3 - u ** 2) * a1)
        # the reference's float32 expression with powf (t2 = t*t exactly rounded; t3 correctly rounded)
        f = np.float32
        t2, t3 = f(t * t), f(np.float64(t) ** 3)
        h00, h10 = f(f(f(2) * t3 - f(3) * t2) + f(1)), f(f(t3 - f(2) * t2) + t)
        h01, h11 = f(f(-2) * t3 + f(3) * t2), f(t3 - t2)
        ref = (((v0 * h00 + (b0 * h10) * f(1)) + v1 * h01) + (a1 * h11) * f(1)).astype(np.float64)
        scale = _ulp32(np.abs(vals).max())
        worst64 = max(worst64, float(np.abs(got - exact).max() / scale))
        worst_ref = max(worst_ref, float(np.abs(got - ref).max() / scale))
    print(f"cubic spline: {worst64:.2f} ulp from float64, {worst_ref:.2f} ulp from the powf form")
    assert worst64 <= 8.0 and worst_ref <= 8.0
    # slerp: within a few ulps of the float64 rotation by the interpolated angle
    sp = sampler(2, N.INTERP_LINEAR)
    worst = 0.0
    for _ in range(400):
        axis = rng.normal(size=3)
        a0, a1_ = rng.uniform(-3, 3, 2)
        q0, q1 = axis_quat(axis, a0), axis_quat(axis, a1_)
        vals = np.concatenate([q0, q1]).astype(np.float32)
        t = float(np.float32(rng.uniform(0.01, 0.99)))
        got = oracle.anim_sample(sp, times, vals, 4, t).astype(np.float64)
        a, b = vals[:4].astype(np.float64), vals[4:].astype(np.float64)
        c = float(a @ b)
        if c < 0:
            b, c = -b, -c
        if c > 1 - 1.1920928955078125e-7:
            continue                                                 # the lerp branch
        ang = np.arccos(c)
        want = (np.sin((1 - t) * ang) * a + np.sin(t * ang) * b) / np.sin(ang)
        worst = max(worst, float(np.abs(got - want).max() / _ulp32(1.0)))
    print(f"slerp: {worst:.2f} ulp of 1.0 from float64")
    assert worst <= 4.0
