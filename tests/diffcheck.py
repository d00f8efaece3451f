"""FP64 gradient checks of the differentiable form (test helper).

The function differentiated is the forward with the rasterization held
fixed (ids and depth are inputs, without a gradient), as in the extension:

    L(pos, attr, xforms) = sum (antialias(interpolate(clip, attr, ids)) - target)^2
    clip[c, i, v] = VP_c @ M_i @ [pos_v, 1]

`analytic` chains the oracle's FP64 gradient with respect to clip and attr
through the two matrix products; `central_differences` evaluates the FP64
loss at theta +- h for chosen parameters.  The loss is evaluated only on the
pixels a parameter can reach (pixels whose triangle uses the vertex or
instance, and their four neighbours), which leaves its differences unchanged.
Like the ids, the antialiasing state of every pair (which edge it crosses,
which pixel it blends into) is held at its value at theta: the stencil then
stays on the smooth branch the analytic gradient differentiates.

Differences are Ridders-extrapolated: central differences from an initial
step of 1e-3, halved per level, combined in a Neville tableau that cancels
the h^2, h^4, ... terms of their error (C. J. F. Ridders, "Accurate
computation of F'(x) and F'(x)F''(x)", Adv. Eng. Software 4 (1982) 75-76).
A plain central difference at 1e-3 would not do: barycentrics vary on the
scale of a triangle (a C4 edge is about 0.016 units), so its truncation
error reaches percent level, far above the 1e-4 being checked.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import oracle


@dataclass
class Problem:
    pos: np.ndarray         # [V, 3] float64
    attr: np.ndarray        # [V, A] float64
    xforms: np.ndarray      # [I, 4, 4] float64
    vps: np.ndarray         # [C, 4, 4] float64
    tris: np.ndarray        # [T, 3] int32
    ids: np.ndarray         # [C, H, W] int32 (fixed)
    depth: np.ndarray       # [C, H, W] float32 (fixed)
    target: np.ndarray      # [C, H, W, A] float64
    aa: bool = True
    at_clip: np.ndarray | None = None   # evaluate at these clip positions instead

    def clip(self, pos=None, xforms=None):
        if pos is None and xforms is None and self.at_clip is not None:
            return self.at_clip
        pos = self.pos if pos is None else pos
        xforms = self.xforms if xforms is None else xforms
        ph = np.concatenate([pos, np.ones((len(pos), 1))], 1)
        A = np.einsum("cab,ibd->ciad", self.vps, xforms)
        return np.einsum("ciad,vd->civa", A, ph)


def clip32(pos, xforms, vps):
    """Float32 clip positions [C, I, V, 4] (computed in float64, rounded)."""
    ph = np.concatenate([np.asarray(pos, np.float64), np.ones((len(pos), 1))], 1)
    A = np.einsum("cab,ibd->ciad", np.asarray(vps, np.float64), np.asarray(xforms, np.float64))
    return np.einsum("ciad,vd->civa", A, ph).astype(np.float32)


def analytic(pr: Problem):
    """(g_pos [V, 3], g_attr [V, A], g_xforms [I, 4, 4]) in float64."""
    clip = pr.clip()
    g_clip, g_attr = oracle.diff_backward(clip, pr.attr, pr.tris, pr.ids, pr.depth, pr.target, pr.aa)
    A = np.einsum("cab,ibd->ciad", pr.vps, pr.xforms)
    g_pos = np.einsum("ciad,civa->vd", A, g_clip)[:, :3]
    ph = np.concatenate([pr.pos, np.ones((len(pr.pos), 1))], 1)
    G = np.einsum("civa,vd->ciad", g_clip, ph)
    g_x = np.einsum("cab,ciad->ibd", pr.vps, G)
    return g_pos, g_attr, g_x


class Reach:
    """Pixels reached by a vertex or an instance: those whose triangle uses
    it, and their four neighbours."""

    def __init__(self, pr: Problem):
        T = len(pr.tris)
        self.shape = pr.ids.shape
        flat = pr.ids.ravel()
        cov = np.flatnonzero(flat > 0)
        tri = (flat[cov] - 1) % T
        order = np.argsort(tri, kind="stable")
        self.px_by_tri = cov[order]
        self.tri_sorted = tri[order]
        self.inst = np.where(flat > 0, (flat - 1) // T, -1)
        v2t = [[] for _ in range(pr.pos.shape[0])]
        for t, (a, b, c) in enumerate(pr.tris.tolist()):
            v2t[a].append(t); v2t[b].append(t); v2t[c].append(t)
        self.v2t = v2t

    def _dilate(self, px):
        C, H, W = self.shape
        x = px % W
        y = (px // W) % H
        nb = [px, px[x > 0] - 1, px[x < W - 1] + 1, px[y > 0] - W, px[y < H - 1] + W]
        return np.unique(np.concatenate(nb))

    def vertex(self, v):
        parts = []
        for t in self.v2t[v]:
            lo, hi = np.searchsorted(self.tri_sorted, [t, t + 1])
            parts.append(self.px_by_tri[lo:hi])
        px = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        return self._dilate(px.astype(np.int64))

    def instance(self, i):
        return self._dilate(np.flatnonzero(self.inst == i).astype(np.int64))


def visible_vertices(pr: Problem):
    T = len(pr.tris)
    t = (pr.ids[pr.ids > 0] - 1) % T
    return np.unique(pr.tris[np.unique(t)].ravel())


def choose_params(pr: Problem, count: int, seed: int, max_xform: int = 24):
    """`count` distinct seeded parameters (kind, index tuple): up to
    `max_xform` of the instances' 12 affine transform entries, the rest split
    between coordinates and attribute channels of vertices that reach a
    pixel."""
    rng = np.random.default_rng(seed)
    vis = visible_vertices(pr)
    I = pr.xforms.shape[0]
    A = pr.attr.shape[1]
    xf = [("xform", (i, r, c)) for i in range(I) for r in range(3) for c in range(4)]
    xf = [xf[j] for j in rng.permutation(len(xf))[:max_xform]]
    n_pos = (count - len(xf)) // 2
    n_attr = count - len(xf) - n_pos
    pos_pool = [(int(v), k) for v in vis for k in range(3)]
    attr_pool = [(int(v), k) for v in vis for k in range(A)]
    pos = [("pos", pos_pool[j]) for j in rng.choice(len(pos_pool), n_pos, replace=False)]
    attr = [("attr", attr_pool[j]) for j in rng.choice(len(attr_pool), n_attr, replace=False)]
    out = []
    for k in range(max(len(pos), len(attr), len(xf))):
        for lst in (pos, attr, xf):
            if k < len(lst):
                out.append(lst[k])
    return out


def ridders(f, h0: float = 1e-3, shrink: float = 2.0, levels: int = 16, rtol: float = 1e-9):
    """(derivative at 0, error estimate) of the scalar function f.  The
    tableau keeps shrinking the step while its error estimate is above
    rtol * |estimate| (when h0 exceeds the distance to a singularity of f,
    the first levels are not yet in the range where the expansion holds).
    No error estimate is taken below the rounding noise of the level's
    difference quotient, 4 eps (|f(h)| + |f(-h)|) / 2h: below it two
    differences can agree by accident."""
    eps = np.finfo(np.float64).eps
    a = np.zeros((levels, levels))
    noise = np.zeros(levels)
    h = h0

    def cd(h):
        fp, fm = f(h), f(-h)
        return (fp - fm) / (2 * h), 4 * eps * (abs(fp) + abs(fm)) / (2 * h)

    a[0, 0], noise[0] = cd(h)
    best, err = a[0, 0], np.inf
    c2 = shrink * shrink
    for i in range(1, levels):
        h /= shrink
        a[0, i], noise[i] = cd(h)
        fac = c2
        for j in range(1, i + 1):
            a[j, i] = (a[j - 1, i] * fac - a[j - 1, i - 1]) / (fac - 1.0)
            fac *= c2
            e = max(abs(a[j, i] - a[j - 1, i]), abs(a[j, i] - a[j - 1, i - 1]), noise[i])
            if e <= err:
                err, best = e, a[j, i]
        if abs(a[i, i] - a[i - 1, i - 1]) >= 2.0 * err and err <= rtol * abs(best):
            break
    return best, err


class _Evaluator:
    """FP64 loss over a pixel set as a function of one parameter; the clip
    positions are updated only where the parameter reaches."""

    def __init__(self, pr: Problem):
        self.pr = pr
        self.clip = pr.clip()
        self.A = np.einsum("cab,ibd->ciad", pr.vps, pr.xforms)
        self.ph = np.concatenate([pr.pos, np.ones((len(pr.pos), 1))], 1)

    def loss(self, kind, idx, s, pix, state, mode):
        pr = self.pr
        clip, attr = self.clip, pr.attr
        if kind == "pos":
            v, k = idx
            p = self.ph[v].copy()
            p[k] += s
            clip = clip.copy()
            clip[:, :, v, :] = np.einsum("ciad,d->cia", self.A, p)
        elif kind == "attr":
            attr = attr.copy()
            attr[idx] += s
        else:
            i = idx[0]
            M = pr.xforms[i].copy()
            M[idx[1:]] += s
            clip = clip.copy()
            clip[:, i] = np.einsum("cab,bd,vd->cva", pr.vps, M, self.ph)
        return oracle.diff_loss_at(clip, attr, pr.tris, pr.ids, pr.depth, pix, pr.target, pr.aa, state, mode)


def central_differences(pr: Problem, params, count: int, h0: float = 1e-3, reach: Reach | None = None):
    """[(kind, index, fd, error estimate)] for the first `count` parameters."""
    reach = reach or Reach(pr)
    ev = _Evaluator(pr)
    out = []
    for kind, idx in params[:count]:
        pix = reach.instance(idx[0]) if kind == "xform" else reach.vertex(idx[0])
        state = np.zeros((pix.size, 4), np.int8)
        ev.loss(kind, idx, 0.0, pix, state, 1)
        d, e = ridders(lambda s: ev.loss(kind, idx, s, pix, state, 2), h0)
        out.append((kind, idx, d, e))
    return out


def pick(grads, kind, idx):
    g_pos, g_attr, g_x = grads
    return {"pos": g_pos, "attr": g_attr, "xform": g_x}[kind][idx]


def relative_errors(fd, grads, rel_floor=1e-6):
    """|g - fd| / max(|fd|, rel_floor * max |g| over the parameter's kind)
    for every checked parameter."""
    scale = {k: float(np.abs(g).max()) for k, g in zip(("pos", "attr", "xform"), grads)}
    return np.array([abs(pick(grads, f[0], f[1]) - f[2]) / max(abs(f[2]), rel_floor * scale[f[0]]) for f in fd])


def reaching_fraction(fd, grads, rel=1e-3):
    """Fraction of the checked parameters whose derivative is above rel
    times the largest gradient entry of their kind."""
    scale = {k: float(np.abs(g).max()) for k, g in zip(("pos", "attr", "xform"), grads)}
    return float(np.mean([abs(f[2]) > rel * scale[f[0]] for f in fd]))
