/*
 * diffref.c - test infrastructure: scalar reference of the differentiable
 * form (vk_gltf_viewer_amd/csrc/diff/diff.hip), used only by tests/ as the
 * checker of the HIP kernels.
 *
 *  ref_diff_raster    the triangle raster in float with the device's rules
 *                     (snap to 1/256 px, exact integer edges, top-left tie,
 *                     double depth-plane setup, float evaluation) -> ids, depth
 *  ref_diff_forward   interpolation + antialiasing + L2 loss in double
 *  ref_diff_backward  the analytic gradient of that loss in double
 *  ref_diff_loss_at   that loss over a list of pixels, optionally with the
 *                     antialiasing state held fixed (finite differences)
 *
 * The interpolation and antialiasing definitions are documented in diff.hip;
 * this file restates them independently in double precision.  The analytic
 * gradient is itself pinned by central finite differences of
 * ref_diff_forward (tests/test_oracle_diff.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int C, I, V, T, H, W, A; } ddims;

static size_t cidx(const ddims* d, int c, int i, int v) { return (((size_t)c * d->I + i) * d->V + v) * 4; }

/* ------------------------------------------------------------------------ */
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

void ref_diff_raster(const float* clip, const int* tris, int C, int I, int V, int T, int H, int W,
                     int* ids, float* depth)
{
    const ddims d = { C, I, V, T, H, W, 0 };
    const size_t npx = (size_t)C * H * W;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (npx ? npx : 1));
    for (size_t p = 0; p < npx; ++p) keys[p] = ~0ull;
    const float hw = (float)W * 0.5f, hh = (float)H * 0.5f;
    for (int c = 0; c < C; ++c)
    for (int i = 0; i < I; ++i)
    for (int t = 0; t < T; ++t) {
        int64_t X[3], Y[3];
        float Z[3];
        int ok = 1;
        for (int k = 0; k < 3 && ok; ++k) {
            const float* p = clip + cidx(&d, c, i, tris[3 * t + k]);
            if (!(p[3] > 0.0f)) { ok = 0; break; }
            const float xf = hw * (p[0] / p[3]) + hw, yf = hh * (p[1] / p[3]) + hh;
            if (!(fabsf(xf) < 1048576.0f && fabsf(yf) < 1048576.0f)) { ok = 0; break; }
            X[k] = (int64_t)rintf(xf * 256.0f);
            Y[k] = (int64_t)rintf(yf * 256.0f);
            Z[k] = p[2] / p[3];
        }
        if (!ok) continue;
        int64_t area = (X[1] - X[0]) * (Y[2] - Y[0]) - (Y[1] - Y[0]) * (X[2] - X[0]);
        if (area == 0) continue;
        if (area < 0) {
            int64_t tx = X[1]; X[1] = X[2]; X[2] = tx;
            tx = Y[1]; Y[1] = Y[2]; Y[2] = tx;
            const float tz = Z[1]; Z[1] = Z[2]; Z[2] = tz;
            area = -area;
        }
        int64_t mnx = X[0], mxx = X[0], mny = Y[0], mxy = Y[0];
        for (int k = 1; k < 3; ++k) {
            if (X[k] < mnx) mnx = X[k]; if (X[k] > mxx) mxx = X[k];
            if (Y[k] < mny) mny = Y[k]; if (Y[k] > mxy) mxy = Y[k];
        }
        int64_t px0 = (mnx - 128 + 255) >> 8, px1 = (mxx - 128) >> 8, py0 = (mny - 128 + 255) >> 8, py1 = (mxy - 128) >> 8;
        if (px0 < 0) px0 = 0;
        if (py0 < 0) py0 = 0;
        if (px1 > W - 1) px1 = W - 1;
        if (py1 > H - 1) py1 = H - 1;
        if (px0 > px1 || py0 > py1) continue;
        int64_t ex[3], ey[3], ax[3], ay[3];
        int tl[3];
        for (int k = 0; k < 3; ++k) {
            const int ka = (k + 1) % 3, kb = (k + 2) % 3;
            ex[k] = X[kb] - X[ka]; ey[k] = Y[kb] - Y[ka]; ax[k] = X[ka]; ay[k] = Y[ka];
            tl[k] = ey[k] < 0 || (ey[k] == 0 && ex[k] > 0);
        }
        const double e1x = (double)(X[1] - X[0]) / 256.0, e1y = (double)(Y[1] - Y[0]) / 256.0;
        const double e2x = (double)(X[2] - X[0]) / 256.0, e2y = (double)(Y[2] - Y[0]) / 256.0;
        const double dz1 = (double)Z[1] - (double)Z[0], dz2 = (double)Z[2] - (double)Z[0];
        const double det = (double)area / 65536.0;
        const float za = (float)((dz1 * e2y - dz2 * e1y) / det), zb = (float)((dz2 * e1x - dz1 * e2x) / det);
        const uint32_t id = (uint32_t)(i * T + t + 1);
        for (int64_t py = py0; py <= py1; ++py) {
            const int64_t sy = py * 256 + 128;
            for (int64_t px = px0; px <= px1; ++px) {
                const int64_t sx = px * 256 + 128;
                int in = 1;
                for (int k = 0; k < 3; ++k) {
                    const int64_t e = ex[k] * (sy - ay[k]) - ey[k] * (sx - ax[k]);
                    in = in && (e > 0 || (e == 0 && tl[k]));
                }
                if (!in) continue;
                float z = Z[0] + (za * ((float)(sx - X[0]) * 0.00390625f) + zb * ((float)(sy - Y[0]) * 0.00390625f));
                if (!(z >= 0.0f && z <= 1.0f)) continue;
                z = z + 0.0f;
                const uint64_t key = ((uint64_t)(~f2u(z)) << 32) | (uint64_t)(0xFFFFFFFEu - id);
                uint64_t* dst = keys + (size_t)c * H * W + (size_t)py * W + px;
                if (key < *dst) *dst = key;
            }
        }
    }
    for (size_t p = 0; p < npx; ++p) {
        if (keys[p] == ~0ull) { ids[p] = 0; depth[p] = 0.0f; continue; }
        ids[p] = (int)(0xFFFFFFFEu - (uint32_t)keys[p]);
        depth[p] = u2f(~(uint32_t)(keys[p] >> 32));
    }
    free(keys);
}

/* ------------------------------------------------------------------------ */
static void cross(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

typedef struct {
    int v[3], inst;
    double h[3][3], r[3][3], l[3], S, lam[3];
} dbary;

static void bary(const double* clip, const int* tris, const ddims* d, int c, int id, double xn, double yn, dbary* b)
{
    const int i = (id - 1) / d->T, t = (id - 1) % d->T;
    b->inst = i;
    for (int k = 0; k < 3; ++k) {
        b->v[k] = tris[3 * t + k];
        const double* p = clip + cidx(d, c, i, b->v[k]);
        b->h[k][0] = p[0]; b->h[k][1] = p[1]; b->h[k][2] = p[3];
    }
    b->S = 0.0;
    for (int k = 0; k < 3; ++k) {
        cross(b->h[(k + 1) % 3], b->h[(k + 2) % 3], b->r[k]);
        b->l[k] = b->r[k][0] * xn + b->r[k][1] * yn + b->r[k][2];
        b->S += b->l[k];
    }
    for (int k = 0; k < 3; ++k) b->lam[k] = b->l[k] / b->S;
}

static double ndcx(int px, const ddims* d) { return (px + 0.5) / (d->W * 0.5) - 1.0; }
static double ndcy(int py, const ddims* d) { return (py + 0.5) / (d->H * 0.5) - 1.0; }

typedef struct {
    size_t f, b;
    int to_front, inst, va, vb, kb;
    double t, alpha, ef, eb;
    double sa[2], sb[2], cf[2], cb[2];
    double pa[4], pb[4];
} dpair;

static void screen(const double* p, const ddims* d, double* s)
{
    s[0] = d->W * 0.5 * (p[0] / p[3]) + d->W * 0.5;
    s[1] = d->H * 0.5 * (p[1] / p[3]) + d->H * 0.5;
}

static double efn(const double* a, const double* b, const double* p)
{
    return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]);
}

/* fkb < 0: choose the crossed edge and the blend direction; else use edge
 * fkb and direction ffront as given (the antialiasing state held fixed) */
static int pair_eval_state(const double* clip, const int* tris, const int* ids, const float* depth, const ddims* d,
                           int c, int px, int py, int dir, int fkb, int ffront, dpair* e)
{
    const int qx = px + (dir == 0), qy = py + (dir == 1);
    if (qx >= d->W || qy >= d->H) return 0;
    const size_t base = (size_t)c * d->H * d->W;
    const size_t ip = base + (size_t)py * d->W + px, iq = base + (size_t)qy * d->W + qx;
    const int idp = ids[ip], idq = ids[iq];
    if (idp == idq) return 0;
    int p_front;
    if (idp == 0) p_front = 0;
    else if (idq == 0) p_front = 1;
    else p_front = depth[ip] >= depth[iq];
    e->f = p_front ? ip : iq;
    e->b = p_front ? iq : ip;
    const int idf = p_front ? idp : idq;
    e->cf[0] = (p_front ? px : qx) + 0.5; e->cf[1] = (p_front ? py : qy) + 0.5;
    e->cb[0] = (p_front ? qx : px) + 0.5; e->cb[1] = (p_front ? qy : py) + 0.5;
    e->inst = (idf - 1) / d->T;
    const int t = (idf - 1) % d->T;
    const double* P[3];
    double S[3][2];
    for (int k = 0; k < 3; ++k) {
        P[k] = clip + cidx(d, c, e->inst, tris[3 * t + k]);
        if (!(P[k][3] > 0.0)) return 0;
        screen(P[k], d, S[k]);
    }
    const double area = efn(S[0], S[1], S[2]);
    if (area == 0.0) return 0;
    const double sgn = area > 0.0 ? 1.0 : -1.0;
    double best = 2.0, bef = 0.0, beb = 0.0;
    int kb = -1;
    if (fkb >= 0) {
        kb = fkb;
        bef = sgn * efn(S[kb], S[(kb + 1) % 3], e->cf);
        beb = sgn * efn(S[kb], S[(kb + 1) % 3], e->cb);
        best = bef / (bef - beb);
        e->to_front = ffront;
    } else {
        for (int k = 0; k < 3; ++k) {
            const double ef = sgn * efn(S[k], S[(k + 1) % 3], e->cf);
            const double eb = sgn * efn(S[k], S[(k + 1) % 3], e->cb);
            if (!(ef >= 0.0 && eb <= 0.0 && ef > eb)) continue;
            const double tt = ef / (ef - eb);
            if (tt < best) { best = tt; kb = k; bef = ef; beb = eb; }
        }
        if (kb < 0) return 0;
        e->to_front = best < 0.5;
    }
    e->kb = kb;
    e->t = best;
    e->alpha = e->to_front ? 0.5 - best : best - 0.5;
    e->va = tris[3 * t + kb];
    e->vb = tris[3 * t + (kb + 1) % 3];
    memcpy(e->pa, P[kb], sizeof e->pa);
    memcpy(e->pb, P[(kb + 1) % 3], sizeof e->pb);
    memcpy(e->sa, S[kb], sizeof e->sa);
    memcpy(e->sb, S[(kb + 1) % 3], sizeof e->sb);
    e->ef = sgn * bef;
    e->eb = sgn * beb;
    return 1;
}

static int pair_eval(const double* clip, const int* tris, const int* ids, const float* depth, const ddims* d, int c,
                     int px, int py, int dir, dpair* e)
{
    return pair_eval_state(clip, tris, ids, depth, d, c, px, py, dir, -1, 0, e);
}

/* colour (A values) of the interpolation at one pixel; 0 on background */
static void interp_px(const double* clip, const double* attr, const int* tris, const int* ids, const ddims* d, int c,
                      int px, int py, double* o)
{
    const size_t p = ((size_t)c * d->H + py) * d->W + px;
    if (ids[p] == 0) { for (int a = 0; a < d->A; ++a) o[a] = 0.0; return; }
    dbary b;
    bary(clip, tris, d, c, ids[p], ndcx(px, d), ndcy(py, d), &b);
    for (int a = 0; a < d->A; ++a)
        o[a] = b.lam[0] * attr[b.v[0] * d->A + a] + b.lam[1] * attr[b.v[1] * d->A + a] + b.lam[2] * attr[b.v[2] * d->A + a];
}

/* colour (C*H*W*A) of the interpolation */
static void interp(const double* clip, const double* attr, const int* tris, const int* ids, const ddims* d, double* col)
{
    for (int c = 0; c < d->C; ++c)
    for (int py = 0; py < d->H; ++py)
    for (int px = 0; px < d->W; ++px)
        interp_px(clip, attr, tris, ids, d, c, px, py, col + (((size_t)c * d->H + py) * d->W + px) * d->A);
}

/* Loss sum (out - target)^2; out (optional) receives the image. */
double ref_diff_forward(const double* clip, const double* attr, int A, const int* tris, int C, int I, int V, int T,
                        int H, int W, const int* ids, const float* depth, const double* target, int aa, double* out)
{
    const ddims d = { C, I, V, T, H, W, A };
    const size_t n = (size_t)C * H * W * A;
    double* col = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* res = (double*)malloc(sizeof(double) * (n ? n : 1));
    interp(clip, attr, tris, ids, &d, col);
    memcpy(res, col, sizeof(double) * n);
    if (aa) {
        for (int c = 0; c < C; ++c)
        for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px)
        for (int dir = 0; dir < 2; ++dir) {
            dpair e;
            if (!pair_eval(clip, tris, ids, depth, &d, c, px, py, dir, &e)) continue;
            const size_t tgt = e.to_front ? e.f : e.b, src = e.to_front ? e.b : e.f;
            for (int a = 0; a < A; ++a) res[tgt * A + a] += e.alpha * (col[src * A + a] - col[tgt * A + a]);
        }
    }
    double loss = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double r = res[k] - (target ? target[k] : 0.0);
        loss += r * r;
    }
    if (out) memcpy(out, res, sizeof(double) * n);
    free(col);
    free(res);
    return loss;
}

/* Gradient of the forward's loss with respect to clip (C*I*V*4) and attr (V*A). */
void ref_diff_backward(const double* clip, const double* attr, int A, const int* tris, int C, int I, int V, int T,
                       int H, int W, const int* ids, const float* depth, const double* target, int aa, double* g_clip,
                       double* g_attr)
{
    const ddims d = { C, I, V, T, H, W, A };
    const size_t n = (size_t)C * H * W * A;
    double* res = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* col = (double*)malloc(sizeof(double) * (n ? n : 1));
    double* gcol = (double*)malloc(sizeof(double) * (n ? n : 1));
    ref_diff_forward(clip, attr, A, tris, C, I, V, T, H, W, ids, depth, target, aa, res);
    interp(clip, attr, tris, ids, &d, col);
    memset(g_clip, 0, sizeof(double) * (size_t)C * I * V * 4);
    memset(g_attr, 0, sizeof(double) * (size_t)V * A);
    /* dL/dout = 2 (out - target); identity part of every output */
    for (size_t k = 0; k < n; ++k) gcol[k] = 2.0 * (res[k] - (target ? target[k] : 0.0));
    if (aa) {
        double* gout = (double*)malloc(sizeof(double) * (n ? n : 1));
        memcpy(gout, gcol, sizeof(double) * n);
        for (int c = 0; c < C; ++c)
        for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px)
        for (int dir = 0; dir < 2; ++dir) {
            dpair e;
            if (!pair_eval(clip, tris, ids, depth, &d, c, px, py, dir, &e)) continue;
            const size_t tgt = e.to_front ? e.f : e.b, src = e.to_front ? e.b : e.f;
            double galpha = 0.0;
            for (int a = 0; a < A; ++a) {
                const double go = gout[tgt * A + a];
                gcol[tgt * A + a] -= e.alpha * go;
                gcol[src * A + a] += e.alpha * go;
                galpha += go * (col[src * A + a] - col[tgt * A + a]);
            }
            const double gt = e.to_front ? -galpha : galpha;
            const double den = e.ef - e.eb;
            const double gef = gt * (-e.eb / (den * den)), geb = gt * (e.ef / (den * den));
            const double* Ps[2] = { e.cf, e.cb };
            const double ge[2] = { gef, geb };
            double gax = 0, gay = 0, gbx = 0, gby = 0;
            for (int k = 0; k < 2; ++k) {
                gax += ge[k] * (e.sb[1] - Ps[k][1]);
                gay += ge[k] * (Ps[k][0] - e.sb[0]);
                gbx += ge[k] * (Ps[k][1] - e.sa[1]);
                gby += ge[k] * (e.sa[0] - Ps[k][0]);
            }
            const double hw = W * 0.5, hh = H * 0.5;
            double* ga = g_clip + cidx(&d, c, e.inst, e.va);
            double* gb = g_clip + cidx(&d, c, e.inst, e.vb);
            ga[0] += gax * hw / e.pa[3];
            ga[1] += gay * hh / e.pa[3];
            ga[3] += -(gax * hw * e.pa[0] + gay * hh * e.pa[1]) / (e.pa[3] * e.pa[3]);
            gb[0] += gbx * hw / e.pb[3];
            gb[1] += gby * hh / e.pb[3];
            gb[3] += -(gbx * hw * e.pb[0] + gby * hh * e.pb[1]) / (e.pb[3] * e.pb[3]);
        }
        free(gout);
    }
    /* interpolation backward */
    for (int c = 0; c < C; ++c)
    for (int py = 0; py < H; ++py)
    for (int px = 0; px < W; ++px) {
        const size_t p = ((size_t)c * H + py) * W + px;
        if (ids[p] == 0) continue;
        const double xn = ndcx(px, &d), yn = ndcy(py, &d);
        dbary b;
        bary(clip, tris, &d, c, ids[p], xn, yn, &b);
        const double* go = gcol + p * A;
        double glam[3] = { 0, 0, 0 };
        for (int a = 0; a < A; ++a)
            for (int k = 0; k < 3; ++k) {
                g_attr[b.v[k] * A + a] += b.lam[k] * go[a];
                glam[k] += go[a] * attr[b.v[k] * A + a];
            }
        const double dot = glam[0] * b.lam[0] + glam[1] * b.lam[1] + glam[2] * b.lam[2];
        double gh[3][3] = { { 0 } };
        for (int k = 0; k < 3; ++k) {
            const double gl = (glam[k] - dot) / b.S;
            const double gr[3] = { gl * xn, gl * yn, gl };
            double a1[3], a2[3];
            cross(b.h[(k + 2) % 3], gr, a1);
            cross(gr, b.h[(k + 1) % 3], a2);
            for (int j = 0; j < 3; ++j) { gh[(k + 1) % 3][j] += a1[j]; gh[(k + 2) % 3][j] += a2[j]; }
        }
        for (int k = 0; k < 3; ++k) {
            double* g = g_clip + cidx(&d, c, b.inst, b.v[k]);
            g[0] += gh[k][0];
            g[1] += gh[k][1];
            g[3] += gh[k][2];
        }
    }
    free(res);
    free(col);
    free(gcol);
}

/* Loss of the forward restricted to the listed output pixels (indices into
 * C*H*W), evaluated per pixel as a gather: the pixel's interpolated colour
 * plus the blend of every neighbour pair whose blend target it is.  Equal to
 * ref_diff_forward's loss over those pixels; a parameter that reaches only
 * these pixels has the same finite differences under both.
 * state (4 bytes per listed pixel, one per pair (left, p), (p, right),
 * (up, p), (p, down)) holds the discrete antialiasing state of each pair:
 * 0 no blend, else 1 + crossed edge + 3 * (blend into the front pixel).
 * mode 1 records it, mode 2 evaluates with the recorded state (the smooth
 * branch the analytic gradient differentiates), mode 0 ignores it. */
static double loss_px(const double* clip, const double* attr, const int* tris, const ddims* d, const int* ids,
                      const float* depth, const double* target, int aa, int64_t p, int8_t* state, int mode)
{
    const int A = d->A, H = d->H, W = d->W;
    double col[8], nb[8], res[8];
    const int c = (int)(p / ((int64_t)H * W));
    const int py = (int)((p / W) % H), px = (int)(p % W);
    interp_px(clip, attr, tris, ids, d, c, px, py, col);
    for (int a = 0; a < A; ++a) res[a] = col[a];
    if (aa) {
        const int qx[4] = { px - 1, px, px, px }, qy[4] = { py, py, py - 1, py }, dir[4] = { 0, 0, 1, 1 };
        for (int k = 0; k < 4; ++k) {
            int8_t* st = state ? state + k : NULL;
            if (qx[k] < 0 || qy[k] < 0) { if (st && mode == 1) *st = 0; continue; }
            dpair e;
            int ok;
            if (mode == 2 && st) {
                ok = *st != 0 &&
                     pair_eval_state(clip, tris, ids, depth, d, c, qx[k], qy[k], dir[k], (*st - 1) % 3, (*st - 1) / 3, &e);
            } else {
                ok = pair_eval(clip, tris, ids, depth, d, c, qx[k], qy[k], dir[k], &e);
                if (st && mode == 1) *st = ok ? (int8_t)(1 + e.kb + 3 * e.to_front) : 0;
            }
            if (!ok) continue;
            const size_t tgt = e.to_front ? e.f : e.b, src = e.to_front ? e.b : e.f;
            if ((int64_t)tgt != p) continue;
            const int sx = (int)(src % W), sy = (int)((src / W) % H);
            interp_px(clip, attr, tris, ids, d, c, sx, sy, nb);
            for (int a = 0; a < A; ++a) res[a] += e.alpha * (nb[a] - col[a]);
        }
    }
    double loss = 0.0;
    for (int a = 0; a < A; ++a) {
        const double r = res[a] - (target ? target[p * A + a] : 0.0);
        loss += r * r;
    }
    return loss;
}

double ref_diff_loss_at(const double* clip, const double* attr, int A, const int* tris, int C, int I, int V, int T,
                        int H, int W, const int* ids, const float* depth, const double* target, int aa,
                        const int64_t* pix, int64_t npix, int8_t* state, int mode)
{
    const ddims d = { C, I, V, T, H, W, A };
    if (A < 1 || A > 8) return NAN;
    /* fixed blocks summed in order: the result does not depend on the thread count */
    enum { kBlk = 1024 };
    const int64_t nb = (npix + kBlk - 1) / kBlk;
    double* part = (double*)calloc((size_t)(nb ? nb : 1), sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; ++b) {
        double acc = 0.0;
        const int64_t end = (b + 1) * kBlk < npix ? (b + 1) * kBlk : npix;
        for (int64_t n = b * kBlk; n < end; ++n)
            acc += loss_px(clip, attr, tris, &d, ids, depth, target, aa, pix[n], state ? state + n * 4 : NULL, mode);
        part[b] = acc;
    }
    double loss = 0.0;
    for (int64_t b = 0; b < nb; ++b) loss += part[b];
    free(part);
    return loss;
}
