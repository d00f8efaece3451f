"""Writes the golden fixtures under tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no tests, golden vectors or fixtures for this path, so the
expected outputs here are computed from the definitions the reference relies
on, independently of the oracle's code:

  camera.npz     glm perspectiveRH_ZO / lookAtRH closed forms, the reverse-Z
                 and Y-flip of camera.cpp:38-48,177-179, Gribb-Hartmann planes
                 with the normalisation and w negation of camera.cpp:70-84, in
                 float64.
  hiz.npz        the Hi-Z reduction of hiz_reduce.comp.glsl:21-31 through the
                 MIN-reduction linear sampler (application.cpp:437-458), from
                 the sampler's definition: texel t of a row carries the tent
                 weight max(0, 1 - |U - (t + 0.5)|) at the unnormalised sample
                 position U, out-of-range texels fold onto the edge
                 (clamp-to-edge), the result is the minimum over texels of
                 non-zero weight.  Even sizes are also a plain 2x2 minimum.
  raster.npz     coverage of small triangles at pixel centres with exact
                 rational arithmetic: strictly inside, or on a top edge
                 (horizontal, third vertex below) or a left edge (third vertex
                 to its right), the tie rule the raster uses.
  srgb.npz       IEC 61966-2-1 sRGB encode (srgb.h.glsl:9-15) in float64 and
                 its 8-bit unorm code.
"""
from __future__ import annotations

import math
import os
from fractions import Fraction

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# camera
def perspective_rh_zo(fovy, aspect, zn, zf):
    f = 1.0 / math.tan(fovy / 2.0)
    m = np.zeros((4, 4))          # m[column][row]
    m[0][0] = f / aspect
    m[1][1] = f
    m[2][2] = zf / (zn - zf)
    m[2][3] = -1.0
    m[3][2] = -(zf * zn) / (zf - zn)
    return m


def look_at_rh(eye, center, up):
    eye, center, up = (np.asarray(v, float) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0][0], m[1][0], m[2][0] = s
    m[0][1], m[1][1], m[2][1] = u
    m[0][2], m[1][2], m[2][2] = -f
    m[3][0] = -s @ eye
    m[3][1] = -u @ eye
    m[3][2] = f @ eye
    return m


def colmajor_mul(a, b):
    # (A B) in column-major storage m[c][r]: to row-major, multiply, back
    return (a.T @ b.T).T


def camera_vp(width, height, eye, target, fovy_deg=60.0, up=(0, 1, 0), zn=0.1, zf=1000.0):
    p = perspective_rh_zo(math.radians(fovy_deg), width / height, zn, zf)
    p[1][1] *= -1.0                                        # camera.cpp:178
    rz = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 0], [0, 0, 1, 1]], float)   # z -> w - z
    vp = colmajor_mul(colmajor_mul(rz, p), look_at_rh(eye, target, up))
    rows = vp.T                                            # rows[r] = row r of the matrix
    planes = []
    for k in range(3):
        for sign in (1.0, -1.0):
            pl = rows[3] + sign * rows[k]
            pl = pl / np.linalg.norm(pl[:3])
            pl[3] = -pl[3]
            planes.append(pl)
    return vp, np.array(planes)


# ---------------------------------------------------------------------------
# Hi-Z
def level_sizes(w, h):
    m = max(w, h)
    n = int(math.floor(math.log2(m)))
    return [(max(1, w >> (k + 1)), max(1, h >> (k + 1))) for k in range(n)]


def tent_taps(dst, src):
    """Per destination index: source texels of non-zero linear weight."""
    taps = []
    for x in range(dst):
        uv = np.float32((np.float32(x) + np.float32(0.5)) / np.float32(dst))
        U = float(np.float32(uv * np.float32(src)))
        sel = set()
        for t in range(int(math.floor(U)) - 2, int(math.floor(U)) + 3):
            if max(0.0, 1.0 - abs(U - (t + 0.5))) > 0.0:
                sel.add(min(max(t, 0), src - 1))
        taps.append(sorted(sel))
    return taps


def hiz_tent(depth):
    h, w = depth.shape
    out, src = [], depth.astype(np.float32)
    for (lw, lh) in level_sizes(w, h):
        tx, ty = tent_taps(lw, src.shape[1]), tent_taps(lh, src.shape[0])
        dst = np.empty((lh, lw), np.float32)
        for y in range(lh):
            for x in range(lw):
                dst[y, x] = src[np.ix_(ty[y], tx[x])].min()
        out.append(dst)
        src = dst
    return out


# ---------------------------------------------------------------------------
# raster coverage (exact)
def covered(tri, px, py):
    P = (Fraction(2 * px + 1, 2), Fraction(2 * py + 1, 2))

    def cross(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])

    orient = cross(*tri)
    if orient == 0:
        return False
    for k in range(3):
        A, B, Cv = tri[k], tri[(k + 1) % 3], tri[(k + 2) % 3]
        s = cross(A, B, P) * (1 if orient > 0 else -1)
        if s > 0:
            continue
        if s < 0:
            return False
        if A[1] == B[1]:
            tie_ok = Cv[1] > A[1]                               # top edge
        else:
            x_at = A[0] + (B[0] - A[0]) * (Cv[1] - A[1]) / (B[1] - A[1])
            tie_ok = Cv[0] > x_at                               # left edge
        if not tie_ok:
            return False
    return True


def raster_cases(rng, W=16, H=12, n=48):
    tris, masks = [], []
    while len(tris) < n:
        # vertex positions on the 1/256-pixel grid, some on pixel centres and
        # on shared horizontal/vertical lines to exercise the tie rule
        pts = []
        for _ in range(3):
            if rng.random() < 0.4:
                x = Fraction(2 * int(rng.integers(0, W)) + 1, 2)
                y = Fraction(2 * int(rng.integers(0, H)) + 1, 2)
            else:
                x = Fraction(int(rng.integers(-256, (W + 1) * 256)), 256)
                y = Fraction(int(rng.integers(-256, (H + 1) * 256)), 256)
            pts.append((x, y))
        if rng.random() < 0.3:
            pts[1] = (pts[1][0], pts[0][1])                     # horizontal edge
        if rng.random() < 0.3:
            pts[2] = (pts[0][0], pts[2][1])                     # vertical edge
        a = (pts[1][0] - pts[0][0]) * (pts[2][1] - pts[0][1]) - (pts[1][1] - pts[0][1]) * (pts[2][0] - pts[0][0])
        if a == 0:
            continue
        m = np.zeros((H, W), np.uint8)
        for py in range(H):
            for px in range(W):
                m[py, px] = covered(pts, px, py)
        tris.append([[float(p[0]), float(p[1])] for p in pts])
        masks.append(m)
    return np.array(tris, np.float64), np.array(masks)


# ---------------------------------------------------------------------------
def srgb_encode(c):
    c = np.asarray(c, np.float64)
    return np.where(c < 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0), 1.0 / 2.4) - 0.055)


def main():
    rng = np.random.default_rng(20240607)
    # camera: the C1 view (1280x720, eye (0, 20, 110) -> origin, 60 deg) and a C2-like 4K street view
    views = [(1280, 720, (0.0, 20.0, 110.0), (0.0, 0.0, 0.0)),
             (3840, 2160, (-378.0, 1.7, 0.0), (-377.0, 1.7, 0.0699268119))]
    vps, planes = zip(*(camera_vp(w, h, e, t) for (w, h, e, t) in views))
    np.savez_compressed(os.path.join(OUT, "camera.npz"),
                        sizes=np.array([(w, h) for (w, h, _, _) in views]),
                        eyes=np.array([e for (_, _, e, _) in views]), targets=np.array([t for (_, _, _, t) in views]),
                        vp=np.array(vps), planes=np.array(planes))

    hiz = {}
    for (w, h) in [(32, 16), (37, 23), (135, 67), (61, 1), (1, 45)]:
        d = rng.random((h, w)).astype(np.float32)
        if w > 4 and h > 4:
            d[rng.random((h, w)) < 0.1] = 0.0
        hiz[f"depth_{w}x{h}"] = d
        hiz[f"pyramid_{w}x{h}"] = np.concatenate([lv.ravel() for lv in hiz_tent(d)])
    np.savez_compressed(os.path.join(OUT, "hiz.npz"), **hiz)

    tris, masks = raster_cases(rng)
    np.savez_compressed(os.path.join(OUT, "raster.npz"), tris=tris, masks=masks, width=16, height=12)

    c = np.concatenate([np.linspace(0.0, 1.0, 2049), rng.random(2047), [0.0031308, 0.00313, 0.0031309, 1e-6, 0.5]])
    c = c.astype(np.float32)
    enc = srgb_encode(c.astype(np.float64))
    np.savez_compressed(os.path.join(OUT, "srgb.npz"), c=c, encoded=enc, code=np.floor(enc * 255.0 + 0.5).astype(np.uint8))
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
