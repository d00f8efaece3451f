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
kept and cached in a JSON file.  The cache key joins the caller's key (scene,
views) with what makes a tuning stale: the device's architecture and compute
units, a hash of the library build and the number of knobs.  A cached tuning
of the wrong length is ignored, and a cached tuning is applied only after one
frame under it passes `valid`.

Two searches: "coordinate" (the default) sweeps one knob at a time over its
candidates with the others held at the best so far, and repeats the sweep
until no knob changes (at most three sweeps; about 11 candidates a sweep for
the default set); "grid" times the full product (48 candidates).

The tuner only changes how work is split, and the frame's results are
identical for every tuning (the parity tests run at several tunings).  A
lower `small_max_extent` sends more triangles to the big-triangle queues;
when they are full, those triangles take the slower direct path (the frame's
overflow bits say so).  A caller passes `valid` (e.g. "the frame's overflow
bits are zero"), and candidates whose frames fail it are not chosen.
"""
from __future__ import annotations

import hashlib
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


def _library_hash() -> str:
    from . import _native
    try:
        with open(_native.HIP_LIB, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()[:16]
    except OSError:
        return "unbuilt"


def cache_key(renderer, key: str) -> str:
    """The caller's key joined with the device identity (architecture and
    compute units, when the renderer reports them), the library build and the
    knob count."""
    ident = getattr(renderer, "device_identity", None)
    dev = ident() if callable(ident) else "unknown-device"
    return f"{key}|{dev}|lib-{_library_hash()}|knobs-{len(KNOBS)}"


def _time(renderer, frame, t, warmup, frames):
    renderer.set_tuning(*t)
    for _ in range(warmup):
        frame()
    renderer.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        frame()
    renderer.sync()
    return 1000.0 * (time.perf_counter() - t0) / frames


def autotune(renderer, frame, key: str, candidates=None, warmup: int = 2, frames: int = 5, use_cache: bool = True,
             log=None, valid=None, search: str = "coordinate"):
    """Times `frame()` (which records one frame on the renderer's stream)
    under candidate tunings and applies the fastest for which `valid()`
    (called after the candidate's frames; default: always true) holds.
    Returns (tuning tuple, {tuning: ms per frame}); rejected candidates are
    listed with an infinite time.  If every candidate is rejected the
    context's tuning is left as it was and that tuning is returned."""
    if search not in ("coordinate", "grid"):
        raise ValueError("search is 'coordinate' or 'grid'")
    path = cache_path()
    ckey = cache_key(renderer, key)
    cache = _load(path) if use_cache else {}
    before = tuple(renderer.get_tuning())
    if use_cache and ckey in cache:
        t = tuple(cache[ckey].get("tuning", ()))
        if len(t) == len(before):
            renderer.set_tuning(*t)
            frame()
            renderer.sync()
            if valid is None or valid():
                return t, {}
            renderer.set_tuning(*before)
        if log:
            log(f"autotune {key}: cached tuning {t} not usable, searching")
    cand = candidates or DEFAULT_CANDIDATES
    # one axis per knob the renderer reports; a knob the candidates leave out keeps its value
    axes = [tuple(cand.get(k, (before[i],))) for i, k in enumerate(KNOBS[:len(before)])]
    times = {}

    def measure(t):
        if t not in times:
            ms = _time(renderer, frame, t, warmup, frames)
            ok = valid is None or valid()
            if log:
                log(f"autotune {key}: {t} {ms:.4f} ms" + ("" if ok else " (rejected)"))
            times[t] = ms if ok else float("inf")
        return times[t]

    if search == "grid":
        for t in itertools.product(*axes):
            measure(t)
    else:
        cur = tuple(before[i] if before[i] in axes[i] else axes[i][0] for i in range(len(axes)))
        for _ in range(3):
            changed = False
            for i, axis in enumerate(axes):
                if len(axis) == 1:
                    cur = cur[:i] + (axis[0],) + cur[i + 1:]
                    continue
                best_v = min(axis, key=lambda v: measure(cur[:i] + (v,) + cur[i + 1:]))
                if best_v != cur[i] and measure(cur[:i] + (best_v,) + cur[i + 1:]) < measure(cur):
                    cur = cur[:i] + (best_v,) + cur[i + 1:]
                    changed = True
            if not changed:
                break
        measure(cur)
    best = min(times, key=times.get)
    if times[best] == float("inf"):
        renderer.set_tuning(*before)
        return before, times
    renderer.set_tuning(*best)
    if use_cache:
        cache[ckey] = {"tuning": list(best), "ms": times[best]}
        _store(path, cache)
    return best, times
