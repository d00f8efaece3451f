"""Benchmark of the GPU-driven frame on one MI355X.

Metric (one JSON line on stdout): ms per frame of cull + Hi-Z + raster +
resolve for the occlusion-heavy 4K city (3840x2160, 2048 noise-displaced
int16-quantized subdivision-4 icospheres + 4096 boxes, 10.53M triangles,
127k meshlet draws), two-pass Hi-Z occlusion culling, barycentric resolve.
Inputs are resident in HBM before the timed region; the camera is static, so
every timed frame is a steady-state two-pass frame.

A step is one frame.  W warm-up frames, then K timed frames bracketed by a
stream synchronization on both sides.  Per-pass durations come from HIP
events recorded on the frame's stream in K further frames after the timed
region.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from vk_gltf_viewer_amd import _native as N  # noqa: E402
from vk_gltf_viewer_amd.renderer import Renderer  # noqa: E402
from vk_gltf_viewer_amd.scene import Scene, Texture, camera_from_view  # noqa: E402

HBM_PEAK_GBS = 8000.0


def pyramid_texels(w, h):
    m, n, total = max(w, h), 0, 0
    while (m >> (n + 1)) != 0:
        total += max(1, w >> (n + 1)) * max(1, h >> (n + 1))
        n += 1
    return total


def visible_triangles(sc, draw_ids):
    """Triangles of the listed meshlet draws."""
    d = sc.draws[np.asarray(draw_ids, dtype=np.int64)]
    mesh_of = sc.primitives["mesh_index"][d["primitive_index"]]
    total = 0
    for mi, m in enumerate(sc.meshes):
        sel = mesh_of == mi
        if sel.any():
            total += int(m.meshlets["triangle_count"][d["meshlet_index"][sel]].astype(np.int64).sum())
    return total


# HBM bytes per frame of each pass, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes over this benchmark, each scaled by the same counter's reading for a
# plain 256 MiB copy kernel (profiles/traffic_c2_4k.json records the commands,
# the kernels of each pass and the per-dispatch values they are summed from).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_c2_4k.json")


def measured_traffic(pass_name, W, H):
    """HBM bytes of one frame's launches of the pass; None without a profile
    of this size or pass."""
    t = profile_json(os.path.basename(TRAFFIC_FILE))
    if not t or (t.get("width"), t.get("height")) != (W, H):
        return None
    total = 0.0
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        kib = t[c].get("per_pass_kib", {}).get(pass_name)
        if kib is None:
            return None
        total += kib * t[c]["factor"] * 1024.0
    return round(total)


def profile_json(name):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def meshlet_bytes(sc, draw_ids):
    """Sum over the listed meshlet draws of 12 B per vertex (position) and 3 B
    per triangle (indices): the raster's algorithmic reads."""
    d = sc.draws[np.asarray(draw_ids, dtype=np.int64)]
    mesh_of = sc.primitives["mesh_index"][d["primitive_index"]]
    total = 0
    for mi, m in enumerate(sc.meshes):
        sel = mesh_of == mi
        if sel.any():
            ml = m.meshlets[d["meshlet_index"][sel]]
            total += int(12 * ml["vertex_count"].astype(np.int64).sum() + 3 * ml["triangle_count"].astype(np.int64).sum())
    return total


def batched_views(steps, warmup, log):
    """The batched form: the 1.31M-triangle instanced-sphere scene rendered to
    64 views of uneven size (4 orthographic cascades at 2048x2048, 6 cube
    faces at 512x512, 54 random perspective views up to 1920x1080) as one
    captured-graph batch frame, against the plain loop of 64 single-view
    frames on the same views.  Two-pass Hi-Z, barycentric resolve; tunables
    chosen by the autotuner on the batch frame; batch and loop timed in
    alternating rounds, medians reported."""
    from vk_gltf_viewer_amd.autotune import autotune
    sc = Scene(3, 3).arrays()
    r = Renderer(0)
    try:
        r.load(sc)
        views = [r.create_view(v.width, v.height) for v in sc.views]
        cams = [camera_from_view(v) for v in sc.views]
        batch = r.create_batch(views)
        flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY
        tuning, _ = autotune(r, lambda: batch.render(cams, flags | N.GRAPH), key="c3-64views", use_cache=False,
                             log=log, valid=lambda: batch.stats().overflow == 0, frames=8)

        def run(fn, n):
            r.sync()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            r.sync()
            return 1000.0 * (time.perf_counter() - t0) / n

        batch_frame = lambda: batch.render(cams, flags | N.GRAPH)

        def loop_frame():
            for v, c in zip(views, cams):
                v.render(c, flags)
        run(batch_frame, warmup)
        run(loop_frame, max(1, warmup // 2))
        # batch and loop timed in alternating rounds (medians), so that clock
        # changes over the run affect both sides alike
        rounds = 5
        ms_b, ms_l = [], []
        for _ in range(rounds):
            ms_b.append(run(batch_frame, max(2, steps // rounds)))
            ms_l.append(run(loop_frame, max(1, steps // (2 * rounds))))
        ms_batch, ms_loop = float(np.median(ms_b)), float(np.median(ms_l))
        st = batch.stats()
        px = sum(v.width * v.height for v in sc.views)
        out = {"views": len(views), "pixels": px, "triangles": Scene(3, 3).triangle_count,
               "ms_per_batch_frame": round(ms_batch, 4), "views_per_s": round(len(views) / ms_batch * 1000.0, 1),
               "loop_ms_per_frame": round(ms_loop, 4), "loop_views_per_s": round(len(views) / ms_loop * 1000.0, 1),
               "speedup_vs_loop": round(ms_loop / ms_batch, 3), "tuning": list(tuning),
               "overflow": int(st.overflow)}
        batch.close()
        return out
    finally:
        r.close()


def uneven_views(steps, warmup, log):
    """The uneven stress configuration: 200 meshes with log-uniform meshlet
    counts (1 to 20k; 39.7M unique triangles) whose source positions are
    fp32, int16-normalized or uint8-normalized, 1-64 instances each (15.2M
    meshlet draws), rendered to 16 views of mixed resolution up to 3840x2160
    as one captured-graph batch frame, against the loop of 16 single-view
    frames.  Two-pass Hi-Z, barycentric resolve; tunables from the
    autotuner on the batch frame."""
    from vk_gltf_viewer_amd.autotune import autotune
    t0 = time.perf_counter()
    scene = Scene(5, 5)
    sc = scene.arrays()
    gen_s = time.perf_counter() - t0
    r = Renderer(0)
    try:
        r.load(sc)
        views = [r.create_view(v.width, v.height) for v in sc.views]
        cams = [camera_from_view(v) for v in sc.views]
        batch = r.create_batch(views)
        flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY
        tuning, _ = autotune(r, lambda: batch.render(cams, flags | N.GRAPH), key="c5-16views", use_cache=False,
                             log=log, warmup=1, frames=3, valid=lambda: batch.stats().overflow == 0)

        def run(fn, n):
            r.sync()
            t = time.perf_counter()
            for _ in range(n):
                fn()
            r.sync()
            return 1000.0 * (time.perf_counter() - t) / n

        batch_frame = lambda: batch.render(cams, flags | N.GRAPH)

        def loop_frame():
            for v, c in zip(views, cams):
                v.render(c, flags)
        run(batch_frame, warmup)
        run(loop_frame, max(1, warmup // 2))
        # batch and loop timed in alternating rounds (medians), so that clock
        # changes over the run affect both sides alike
        rounds = 5
        ms_b, ms_l = [], []
        for _ in range(rounds):
            ms_b.append(run(batch_frame, max(2, steps // rounds)))
            ms_l.append(run(loop_frame, max(1, steps // (2 * rounds))))
        ms_batch, ms_loop = float(np.median(ms_b)), float(np.median(ms_l))
        st = batch.stats()
        # backward to the per-instance transforms: every view's antialiased
        # linear image and its backward (an L2 loss against a zero target,
        # d loss / d image = 2 image), accumulated into one gradient
        import torch
        batch_frame()
        r.sync()
        T = sc.transforms.shape[0]
        g = torch.zeros((T, 16), dtype=torch.float64, device="cuda")
        imgs = [torch.empty((v.height, v.width, 3), dtype=torch.float32, device="cuda") for v in views]

        def grad_step():
            g.zero_()
            for v, c, im in zip(views, cams, imgs):
                v.aa_image(c, im)
                v.aa_backward(c, im, g)
        # a stream of its own (not torch's default, which the library would fence)
        gs = torch.cuda.Stream()
        with torch.cuda.stream(gs):
            for _ in range(max(1, warmup)):
                grad_step()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(gs)
            for _ in range(steps):
                grad_step()
            e1.record(gs)
            torch.cuda.synchronize()
        ms_grad = e0.elapsed_time(e1) / steps
        grad_live = int((g.abs().amax(1) > 0).sum().item())
        vis_draws = [int(len(v.visible(0)) + len(v.visible(1))) for v in views]
        vis_tris = sum(visible_triangles(sc, np.concatenate([v.visible(0), v.visible(1)])) for v in views)
        px = sum(v.width * v.height for v in sc.views)
        out = {"views": len(views), "pixels": px, "meshes": len(sc.meshes), "draws": len(sc.draws),
               "triangles": scene.triangle_count, "visible_draws": sum(vis_draws), "visible_triangles": vis_tris,
               "ms_per_batch_frame": round(ms_batch, 4), "views_per_s": round(len(views) / ms_batch * 1000.0, 1),
               "visible_triangles_per_s": round(vis_tris / (ms_batch * 1e-3), 1),
               "loop_ms_per_frame": round(ms_loop, 4), "speedup_vs_loop": round(ms_loop / ms_batch, 3),
               "tuning": list(tuning), "overflow": int(st.overflow), "scene_generation_s": round(gen_s, 2),
               "transform_backward_ms": round(ms_grad, 4), "transforms": T, "transforms_with_gradient": grad_live}
        batch.close()
        return out
    finally:
        r.close()


def differentiable(steps, warmup, precisions=("mixed", "float32")):
    """The differentiable form: the frequency-71 geodesic icosphere (50,412
    vertices, 100,820 triangles) as 8 instances, 8 ring cameras at 512x512, L2 loss of
    the antialiased image against a target rendered from a perturbed mesh.
    A step is one optimisation step's GPU work: clip positions, raster,
    interpolation, antialiasing, loss, and the backward to positions, albedo
    and the 8 instance transforms.  Timed with events on torch's current
    stream, which the extension launches on."""
    import torch
    from vk_gltf_viewer_amd import diff as D
    s = D.c4_scene()
    tp, ta = D.c4_target(s)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    tris, vp = t(s.tris), t(s.view_projs)
    out = {"vertices": int(s.pos.shape[0]), "triangles": int(s.tris.shape[0]), "instances": int(s.xforms.shape[0]),
           "cameras": int(s.view_projs.shape[0]), "size": [s.height, s.width]}
    for prec in precisions:
        with torch.no_grad():
            target, _, _ = D.render(t(tp), t(ta), tris, t(s.xforms), vp, s.height, s.width, precision=prec)
        pos, alb, xf = (t(a).requires_grad_() for a in (s.pos, s.albedo, s.xforms))

        def step():
            img, _, _ = D.render(pos, alb, tris, xf, vp, s.height, s.width, precision=prec)
            loss = ((img - target) ** 2).sum()
            loss.backward()
            return loss

        for _ in range(max(1, warmup)):
            step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        npx = out["cameras"] * s.height * s.width
        out[prec] = {"ms_per_step": round(ms, 4), "mpix_per_s": round(npx / (ms * 1e-3) / 1e6, 1)}
    return out


def synthetic_hierarchy(nodes, seed=0, animated_frac=0.5):
    """A random forest of `nodes` glTF nodes (every node a mesh node), half of
    them driven by 1-3 channels of random path and interpolation with 2-16
    keyframes: the input of the device transform update."""
    from vk_gltf_viewer_amd.scene import ANIM_CHANNEL_DTYPE, ANIM_SAMPLER_DTYPE, NODE_DTYPE, Hierarchy
    rng = np.random.default_rng(seed)
    roots = max(1, nodes // 64)
    parent = np.full(nodes, -1, np.int64)
    parent[roots:] = (rng.random(nodes - roots) * np.arange(roots, nodes)).astype(np.int64)
    order = np.argsort(parent, kind="stable")                 # children grouped by parent, in index order
    kids = order[parent[order] >= 0]
    counts = np.bincount(parent[parent >= 0], minlength=nodes)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    nd = np.zeros(nodes, NODE_DTYPE)
    nd["mesh"] = 0
    nd["first_child"], nd["child_count"] = first, counts
    nd["translation"] = rng.normal(0, 1, (nodes, 3))
    q = rng.normal(size=(nodes, 4))
    nd["rotation"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    nd["scale"] = rng.uniform(0.8, 1.2, (nodes, 3))
    anim = np.nonzero(rng.random(nodes) < animated_frac)[0]
    per = rng.integers(1, 4, len(anim))
    ch_node = np.repeat(anim, per)
    n = len(ch_node)
    path = rng.integers(0, 3, n)
    interp = rng.integers(0, 3, n)
    keys = rng.integers(2, 17, n)
    outs = keys * np.where(interp == 2, 3, 1)
    comps = np.where(path == 1, 4, 3)
    t_off = np.concatenate([[0], np.cumsum(keys)[:-1]])
    v_off = np.concatenate([[0], np.cumsum(outs * comps)[:-1]])
    times = np.sort(rng.uniform(0, 4, int(keys.sum())).astype(np.float32))
    # ascending keys within each sampler: split one sorted pool, then re-sort per segment
    times = np.concatenate([np.sort(times[o:o + k]) for o, k in zip(t_off, keys)]).astype(np.float32)
    values = rng.normal(0, 1, int((outs * comps).sum())).astype(np.float32)
    sp = np.zeros(n, ANIM_SAMPLER_DTYPE)
    sp["input_offset"], sp["input_count"], sp["output_offset"], sp["interpolation"] = t_off, keys, v_off, interp
    ch = np.zeros(n, ANIM_CHANNEL_DTYPE)
    ch["node"], ch["sampler"], ch["path"] = ch_node, np.arange(n), path
    return Hierarchy(nd, kids.astype(np.uint32), np.arange(roots, dtype=np.uint32), sp, times, values, ch)


def widened_rows(sc, cam, W, H, steps, warmup, log):
    """The rows beyond the frame: the textured resolve and motion vectors on
    the same 4K workload, and the device transform update of a 262144-node
    animated hierarchy."""
    out = {}
    sc2 = dataclasses.replace(sc)
    ntex, tsz = 16, 1024
    rng = np.random.default_rng(7)
    sc2.textures = [Texture(rng.integers(0, 256, (tsz, tsz, 4), dtype=np.uint8), N.WRAP_REPEAT, N.WRAP_REPEAT)
                    for _ in range(ntex)]
    sc2.materials = sc.materials.copy()
    sc2.materials["albedo_index"] = np.arange(len(sc2.materials)) % ntex
    sc2.meshes = []
    for m in sc.meshes:
        v = m.vertices.copy()
        p = v["position"]
        v["uv"] = np.stack([p[:, 0] * 0.5, p[:, 2] * 0.5 + p[:, 1] * 0.25], 1).astype(np.float16).view(np.uint16)
        sc2.meshes.append(dataclasses.replace(m, vertices=v))
    r = Renderer(0)
    r.load(sc2)
    view = r.create_view(W, H)
    base = N.CULL_FRUSTUM | N.CULL_TWO_PASS
    npx = W * H
    res = {}
    for name, fl in (("bary", base | N.RESOLVE_BARY), ("textured", base | N.RESOLVE_TEXTURED),
                     ("textured_motion", base | N.RESOLVE_TEXTURED | N.MOTION_VECTORS)):
        for _ in range(max(1, warmup)):
            view.render(cam, fl)
        r.sync()
        view.stats()
        for _ in range(steps):
            view.render(cam, fl | N.TIMING)
        r.sync()
        st = view.stats()
        vis_tris = visible_triangles(sc2, np.concatenate([view.visible(0), view.visible(1)]))
        # algorithmic bytes: key 8 + colour 4 per pixel, 39 B per visible triangle;
        # textured: + 16 B (2x2 RGBA8 texels) per pixel + 12 B (3 float16 uv pairs)
        # per visible triangle; motion: + 8 B (float2) per pixel
        alg = 12.0 * npx + 39.0 * vis_tris
        if name != "bary":
            alg += 16.0 * npx + 12.0 * vis_tris
        if name == "textured_motion":
            alg += 8.0 * npx
        res[name] = {"frame_ms": round(st.ms_total, 4), "resolve_ms": round(st.ms_resolve, 4),
                     "resolve_alg_gbs": round(alg / (st.ms_resolve * 1e-3) / 1e9, 1)}
    out["resolve_4k"] = res
    view.close()
    r.close()
    # device transform update
    h = synthetic_hierarchy(1 << 18)
    r = Renderer(0)
    r.set_hierarchy(h)
    for i in range(max(1, warmup)):
        r.update_transforms(0.1 * i)
    r.sync()
    t0 = time.perf_counter()
    for i in range(steps):
        r.update_transforms(0.013 * i)
    r.sync()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    out["transform_update"] = {"nodes": len(h.nodes), "channels": len(h.channels), "mesh_nodes": len(h.nodes),
                               "ms_per_update": round(ms, 4),
                               "mesh_nodes_per_s": round(len(h.nodes) / (ms * 1e-3), 1)}
    r.close()
    if log:
        log(f"widened rows: {json.dumps(out)}")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, cam, W, H, flags, budget_s=12.0, max_frames=8):
    """The oracle (the scalar C restatement, its frame loops parallel over all
    the host's threads with OpenMP) on frames of the same workload: two
    warm-up frames (the first fills the history, the second still rasterizes
    every draw it listed), then steady-state frames within the budget."""
    import oracle
    threads = oracle.threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))
    rs = oracle.Scene(sc)
    hist = np.zeros(len(sc.draws), np.uint8)
    for _ in range(2):
        oracle.frame_two_pass(rs, cam, W, H, flags, hist)
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < max_frames and time.perf_counter() < t_end:
        t0 = time.perf_counter()
        oracle.frame_two_pass(rs, cam, W, H, flags, hist)
        times.append(time.perf_counter() - t0)
    return 1000.0 * float(np.median(times)), len(times), threads


def timed_frames(r, view, cam, flags, steps, warmup):
    """Wall ms per frame over `steps` frames after `warmup` (stream
    synchronizations on both sides), then the per-pass event timings of
    `steps` further frames, then the raster counts of one VKGV_STATS frame."""
    for _ in range(max(1, warmup)):
        view.render(cam, flags)
    r.sync()
    view.stats()                                  # drop warm-up timings
    t0 = time.perf_counter()
    for _ in range(steps):
        view.render(cam, flags)
    r.sync()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    for _ in range(steps):
        view.render(cam, flags | N.TIMING)
    r.sync()
    st = view.stats()
    view.render(cam, flags | N.STATS)
    counts = view.stats()
    return ms, st, counts


def pass_table(sc, view, st, counts, W, H):
    """Per-pass ms and algorithmic bytes (per unit x units):
      cull_early  36 B per draw (12 B record + 24 B AABB) + 64 B per instance
                  transform + 4 B per visible draw id written, + 80 B per
                  transform of the frame's VP M table written
      cull_late   the same + 16 B per Hi-Z probe (2x2 fp32; at most one per draw)
      raster      12 B per vertex + 3 B per triangle of every rasterized meshlet
                  + 8 B per fragment (one 64-bit key update), both passes
      hiz         4 B per depth pixel read + 4 B per pyramid texel written
      resolve     8 B key read + 4 B RGBA8 written + 8 B clear key written
                  into the next frame's visibility buffer per pixel + 39 B
                  per visible triangle (3 x 12 B positions + 3 B indices)
    (The frame has no clear pass: the resolve clears the next frame's keys.)"""
    npx = W * H
    vis = np.concatenate([view.visible(0), view.visible(1)])
    n_draw, n_inst = len(sc.draws), len(sc.transforms)
    ms = {"cull_early": st.ms_cull, "cull_late": st.ms_cull_late,
          "raster": st.ms_raster + st.ms_raster_late, "hiz": st.ms_hiz, "resolve": st.ms_resolve}
    alg = {"cull_early": 36.0 * n_draw + 144.0 * n_inst + 4.0 * st.visible_early,
           "cull_late": 36.0 * n_draw + 64.0 * n_inst + 16.0 * n_draw + 4.0 * st.visible_late,
           "raster": float(meshlet_bytes(sc, vis)) + 8.0 * counts.fragments,
           "hiz": 4.0 * npx + 4.0 * pyramid_texels(W, H),
           "resolve": 20.0 * npx + 39.0 * visible_triangles(sc, vis)}
    return ms, alg, vis


def raster_object(r, st, counts, ms_frame, log):
    """Raster throughput against its ceilings: rasterized triangles, fragments
    and 64-bit atomics per second over the raster passes' event time, the
    atomics against the device's measured atomicMin rate, and the raster
    kernels' VALU busy from the committed counter profile."""
    ms_r = st.ms_raster + st.ms_raster_late
    peak_c = r.measure_atomic_rate(False)
    peak_s = r.measure_atomic_rate(True)
    a_per_s = counts.fragment_atomics / (ms_r * 1e-3)
    out = {"ms": round(ms_r, 4), "triangles": int(counts.raster_triangles), "fragments": int(counts.fragments),
           "atomics": int(counts.fragment_atomics), "tile_stores": int(counts.fragment_stores),
           "triangles_per_s": round(counts.raster_triangles / (ms_r * 1e-3), 1),
           "fragments_per_s": round(counts.fragments / (ms_r * 1e-3), 1),
           "atomics_per_s": round(a_per_s, 1),
           "atomic_peak_coalesced_per_s": round(peak_c, 1), "atomic_peak_scattered_per_s": round(peak_s, 1),
           "atomic_frac": round(a_per_s / peak_c, 4),
           "rasterized_triangles_per_s_frame": round(counts.raster_triangles / (ms_frame * 1e-3), 1)}
    pmc = profile_json("raster_pmc_c2_4k.json")
    if pmc:
        out["valu_busy"] = pmc.get("valu_busy")
        out["valu_busy_source"] = "profiles/raster_pmc_c2_4k.json"
    if log:
        log(f"raster: {json.dumps(out)}")
    return out


def cold_frames(view, cam, flags, steps):
    """The first frame of a view (no history, pyramid of an empty early
    pass): the early pass draws nothing and the late pass every draw that
    survives the frustum.  Event time per frame."""
    tot, late = 0.0, 0
    for _ in range(steps):
        view.reset_history()
        view.render(cam, flags | N.TIMING)
        st = view.stats()
        tot += st.ms_total
        late = st.visible_late
    return {"ms_per_frame_events": round(tot / steps, 4), "visible_early": 0, "visible_late": int(late)}


def moving_frames(r, view, vd, flags, steps, warmup, step_m=1.5):
    """Two-pass frames with the camera moving `step_m` metres per frame down
    the street (previous-frame matrices from the frame before): the late pass
    draws what the motion uncovers."""
    from vk_gltf_viewer_amd.scene import camera_advance
    vd = N.ViewDesc.from_buffer_copy(vd)
    cam = camera_from_view(vd)
    view.reset_history()
    cams = []
    for _ in range(warmup + steps):
        vd.eye[0] += step_m
        vd.target[0] += step_m
        cam = camera_advance(cam, vd)
        cams.append(cam)
    for c in cams[:warmup]:
        view.render(c, flags)
    r.sync()
    late = 0
    t0 = time.perf_counter()
    for c in cams[warmup:]:
        view.render(c, flags)
    r.sync()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    st = view.stats()
    late = st.visible_late
    return {"ms_per_frame": round(ms, 4), "metres_per_frame": step_m, "frames": steps,
            "last_visible_early": int(st.visible_early), "last_visible_late": int(late)}


def frame_workload(config, variant, steps, warmup, log, with_cpu, extra=True, tuning=(0, 0, 0)):
    """The occlusion-heavy 4K city frame: steady-state two-pass frames
    (value), per-pass rooflines, raster counts, and the cold and moving
    frames of the same scene."""
    t0 = time.perf_counter()
    scene = Scene(config, config, variant)
    sc = scene.arrays()
    tris = scene.triangle_count
    vd = sc.views[0]
    W, H = vd.width, vd.height
    cam = camera_from_view(vd)
    gen_s = time.perf_counter() - t0
    r = Renderer(0)
    try:
        if any(tuning):   # raster tunables given on the command line (0: the library default)
            cur = r.get_tuning()
            r.set_tuning(*[t or c for t, c in zip(tuning, cur[:3])], cur[3])
        r.load(sc)
        view = r.create_view(W, H)
        flags = N.CULL_FRUSTUM | N.CULL_TWO_PASS | N.RESOLVE_BARY
        ms_frame, st, counts = timed_frames(r, view, cam, flags, steps, warmup)
        if st.overflow:
            print("bench.py: WARNING a raster queue was full (direct path taken)", file=sys.stderr)
        ms, alg, vis = pass_table(sc, view, st, counts, W, H)
        gbs = {k: round(alg[k] / (ms[k] * 1e-3) / 1e9, 1) for k in alg if ms[k] > 0}
        dom = max(ms, key=lambda k: ms[k])
        ach = alg[dom] / (ms[dom] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, W, H),
                    "alg_bytes": round(alg[dom]), "ms": round(ms[dom], 4)}
        out = {"scene": {"config": config, "variant": variant, "width": W, "height": H, "triangles": tris,
                         "draws": len(sc.draws), "visible_early": int(st.visible_early),
                         "visible_late": int(st.visible_late), "generation_s": round(gen_s, 2)},
               "ms_frame": ms_frame, "events_ms_total": round(st.ms_total, 4),
               "passes_ms": {k: round(v, 4) for k, v in ms.items()}, "passes_alg_gbs": gbs,
               "passes_traffic": {k: measured_traffic(k, W, H) for k in ms},
               # measured HBM bytes over algorithmic bytes: above 1, bytes moved that the
               # pass's algorithm does not need (Hi-Z: 8-byte keys read for their 4-byte depth)
               "passes_traffic_ratio": {k: (round(measured_traffic(k, W, H) / alg[k], 3)
                                            if measured_traffic(k, W, H) and alg[k] else None) for k in ms},
               "roofline": roofline, "raster": raster_object(r, st, counts, ms_frame, log), "sc": sc, "cam": cam}
        if extra:
            out["cold"] = cold_frames(view, cam, flags, max(3, steps // 4))
            out["moving"] = moving_frames(r, view, vd, flags, steps, warmup)
        if log:
            log(f"frame {config}.{variant}: tuning {r.get_tuning()}; stats: big tris {st.big_triangles} "
                f"tile pairs {st.big_tiles} max/tile {st.max_tile_tris} clipped {st.clipped_triangles}; "
                f"passes {json.dumps(out['passes_ms'])}")
        view.close()
        if with_cpu:
            cms, nframes, threads = cpu_baseline(sc, cam, W, H, flags)
            out["cpu"] = {"value": round(cms, 1), "unit": "ms/frame", "cores": threads, "kind": "port",
                          "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                          "sample": f"oracle (C restatement, OpenMP over {threads} threads) steady-state two-pass "
                                    f"frames of the same {W}x{H} workload, median of {nframes}"}
        return out
    finally:
        r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--small-max", type=int, default=0, help="raster tunable (0: library default)")
    ap.add_argument("--blocks-per-cu", type=int, default=0, help="raster tunable (0: library default)")
    ap.add_argument("--tile-item-tris", type=int, default=0, help="raster tunable (0: library default)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-view batched-vs-loop measurement")
    ap.add_argument("--batch-only", action="store_true", help="only the 64-view batched-vs-loop measurement")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-form measurement")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 16-view uneven stress measurement")
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the literal 83.9M-triangle city (4096 level-5 spheres)")
    ap.add_argument("--uneven-only", action="store_true", help="only the 16-view uneven stress measurement")
    ap.add_argument("--diff-only", action="store_true", help="only the differentiable-form measurement")
    ap.add_argument("--widened-only", action="store_true",
                    help="only the textured-resolve / motion-vector / transform-update measurements")
    ap.add_argument("--no-widened", action="store_true",
                    help="skip the textured-resolve / motion-vector / transform-update measurements")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 256 MiB device-to-device copies (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": a.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(views shard one per GPU, see the design notes)", file=sys.stderr)
        sys.exit(2)

    log = (lambda m: print(m, file=sys.stderr)) if args.verbose else None
    if args.batch_only:
        print(json.dumps({"batched": batched_views(args.steps, args.warmup, log)}))
        return
    if args.uneven_only:
        print(json.dumps({"uneven": uneven_views(args.steps, args.warmup, log)}))
        return
    if args.diff_only:
        print(json.dumps({"differentiable": differentiable(args.steps, args.warmup)}))
        return

    if args.widened_only:
        sc = Scene(args.config, args.config, args.variant).arrays()
        vd = sc.views[0]
        print(json.dumps({"widened": widened_rows(sc, camera_from_view(vd), vd.width, vd.height, args.steps,
                                                  args.warmup, log)}))
        return

    tuning = (args.small_max, args.blocks_per_cu, args.tile_item_tris)
    f = frame_workload(args.config, args.variant, args.steps, args.warmup, log, not args.no_cpu_baseline,
                       tuning=tuning)
    sc, cam = f.pop("sc"), f.pop("cam")
    scn = f["scene"]
    W, H, tris = scn["width"], scn["height"], scn["triangles"]
    out = {
        "metric": "ms/frame (cull+HiZ+raster+resolve) 4K 10M tris; batched views/s vs plain loop",
        "value": round(f["ms_frame"], 4), "unit": "ms/frame", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f["ms_frame"], 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded generator)",
        "config": {"workload": f"occlusion-heavy city config {args.config}.{args.variant}: {W}x{H}, "
                               f"{scn['draws']} meshlet draws, {tris} triangles, two-pass Hi-Z, barycentric resolve, "
                               f"static camera (steady state)",
                   "width": W, "height": H, "triangles": tris, "draws": scn["draws"],
                   "visible_early": scn["visible_early"], "visible_late": scn["visible_late"],
                   "parallelism": "single GPU"},
        "passes_ms": f["passes_ms"], "passes_alg_gbs": f["passes_alg_gbs"], "passes_traffic": f["passes_traffic"],
        "passes_traffic_ratio": f["passes_traffic_ratio"],
        "gpu_frame_ms_events": f["events_ms_total"],
        "rasterized_triangles_per_s": f["raster"]["rasterized_triangles_per_s_frame"],
        "roofline": f["roofline"],
        "raster": f["raster"],
        "cpu_baseline": f.get("cpu"),
        "c2_cold": f.get("cold"),
        "c2_moving": f.get("moving"),
    }
    if not args.no_literal:
        g = frame_workload(2, 1, min(args.steps, 10), min(args.warmup, 3), log, False)
        g.pop("sc"), g.pop("cam")
        out["c2_literal"] = {"scene": g["scene"], "ms_per_frame": round(g["ms_frame"], 4),
                             "passes_ms": g["passes_ms"], "raster": {k: g["raster"][k] for k in
                                                                     ("triangles", "fragments", "triangles_per_s",
                                                                      "fragments_per_s", "atomics_per_s")},
                             "cold": g["cold"], "moving": g["moving"]}
    if not args.no_widened:
        out["widened"] = widened_rows(sc, cam, W, H, min(args.steps, 20), args.warmup, log)
    if not args.no_batch:
        out["batched"] = batched_views(min(args.steps, 20), args.warmup, log)
    if not args.no_uneven:
        out["uneven"] = uneven_views(min(args.steps, 10), min(args.warmup, 2), log)
    if not args.no_diff:
        out["differentiable"] = differentiable(min(args.steps, 20), args.warmup)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
