"""Autotuner for the raster tunables of a context (vkgv_set_tuning).

The knobs:
- `small_max_extent` sets the bin boundary. Triangles whose bounding box is
  smaller stay in their meshlet's wave; larger ones go to the binned tile pass.
- `raster_blocks_per_cu` sets the persistent raster grid, i.e. the
  wave-to-meshlet mapping.
- `tile_item_tris` is the tile pass's work-item size.
- `resolve_cache_slots` is how many rasterized meshlets per view leave their
  triangles for the barycentric resolve: the raster writes 48 bytes per
  triangle of a cached meshlet, the resolve recomputes the others from the
  scene buffers.  Many more rasterized triangles than pixels favour 0.

A candidate is timed on the caller's own frame: a function that records one
frame (a batch or a view) on the context's stream. The fastest candidate is
kept and cached in a JSON file under a key the caller chooses (device, scene,
views).

The tuner only changes how work is split, and the frame's results are
identical for every tuning (the parity tests run at several tunings).  A
lower `small_max_extent` sends more triangles to the big-triangle queues;
when they are full, those triangles take the slower direct path (the frame's
overflow bits say so).  A caller passes `valid` (e.g. "the frame's overflow
bits are zero"), and candidates whose frames fail it are not chosen.
"""
from __future__ import annotations

import itertools
import json
import os
import time

DEFAULT_CANDIDATES = {
    "small_max_extent": (4, 8, 16, 32),
    "raster_blocks_per_cu": (2, 3),
    "tile_item_tris": (8, 16, 32),
    "resolve_cache_slots": (0, 8192),
}
KNOBS = ("small_max_extent", "raster_blocks_per_cu", "tile_item_tris", "resolve_cache_slots")


def cache_path() -> str:
    return os.environ.get("VKGV_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "vk_gltf_viewer_amd", "tune.json"))


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass        # a read-only home is not an error: the result is still applied


def autotune(renderer, frame, key: str, candidates=None, warmup: int = 2, frames: int = 5, use_cache: bool = True,
             log=None, valid=None):
    """Times `frame()` (which records one frame on the renderer's stream)
    under each candidate tuning and applies the fastest for which `valid()`
    (called after the candidate's frames; default: always true) holds.
    Returns (tuning tuple, {tuning: ms per frame}); rejected candidates are
    listed with an infinite time.  If every candidate is rejected the
    context's tuning is left as it was and that tuning is returned."""
    path = cache_path()
    cache = _load(path) if use_cache else {}
    if use_cache and key in cache:
        t = tuple(cache[key]["tuning"])
        renderer.set_tuning(*t)
        return t, {}
    cand = candidates or DEFAULT_CANDIDATES
    before = tuple(renderer.get_tuning())
    # one axis per knob the renderer reports; a knob the candidates leave out keeps its value
    axes = [cand.get(k, (before[i],)) for i, k in enumerate(KNOBS[:len(before)])]
    times = {}
    for t in itertools.product(*axes):
        renderer.set_tuning(*t)
        for _ in range(warmup):
            frame()
        renderer.sync()
        t0 = time.perf_counter()
        for _ in range(frames):
            frame()
        renderer.sync()
        times[t] = 1000.0 * (time.perf_counter() - t0) / frames
        ok = valid is None or valid()
        if log:
            log(f"autotune {key}: {t} {times[t]:.4f} ms" + ("" if ok else " (rejected)"))
        if not ok:
            times[t] = float("inf")
    best = min(times, key=times.get)
    if times[best] == float("inf"):
        renderer.set_tuning(*before)
        return before, times
    renderer.set_tuning(*best)
    if use_cache:
        cache[key] = {"tuning": list(best), "ms": times[best]}
        _store(path, cache)
    return best, times
