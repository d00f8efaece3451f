// Differentiable form of the frame: triangle raster to an id/depth buffer,
// perspective-correct attribute interpolation and analytic edge antialiasing,
// each with a hand-written backward that recomputes barycentrics and edge
// crossings from the id buffer instead of storing per-pixel attributes.
//
// Conventions (shared with the torch bindings and the FP64 reference):
//  clip      [C, I, V, 4]  clip-space position of vertex v of instance i seen
//                          by camera c (reverse-Z, as the frame's cameras)
//  tris      [T, 3] int32  vertex indices, one mesh shared by the instances
//  ids       [C, H, W] int32  0 = background, else i * T + t + 1
//  depth     [C, H, W] f32    z / w at the pixel centre (larger = nearer)
//  pixel (x, y) has its centre at NDC ((x + 0.5) / (W/2) - 1, (y + 0.5) / (H/2) - 1).
//
// Interpolation: with h_k = (x, y, w) of vertex k's clip position,
// r_k = h_{k+1} x h_{k+2}, l_k = r_k . (xn, yn, 1) and lambda_k = l_k / sum(l),
// colour = sum_k lambda_k attr_k.  The kernels evaluate l_k in the equivalent
// pixel-relative form l_k = u_{k+1} x u_{k+2} (2-D cross product) with
// u_k = (x_k - xn w_k, y_k - yn w_k): subtracting the pixel's ray before the
// product keeps float32 accurate for triangles much smaller than their
// distance, where the 3-D cross product cancels.
//
// Antialiasing: for every horizontally or vertically adjacent pixel pair with
// different ids, the front pixel f (the non-background one, else the one with
// the larger depth) keeps its triangle T; the segment from f's centre to the
// other centre b leaves T through the edge of T whose edge function changes
// sign first, at parameter t in [0, 1] (an edge through a centre counts, so
// the output stays continuous when an edge crosses a centre).  The part of the pixel on the far side of the
// edge, along the pair's axis, takes the other pixel's colour:
//   t < 1/2: out_f += (1/2 - t) (c_b - c_f);   t > 1/2: out_b += (t - 1/2) (c_f - c_b).
// The output is continuous in the vertex positions; the backward gives the
// gradient of the colours and, through t, of the edge's clip positions.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "diff.hpp"

namespace vkgv_diff {
namespace {

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t float_bits(float f) { return __float_as_uint(f); }

__device__ __forceinline__ unsigned long long make_key(float z, uint32_t id)
{
    return ((unsigned long long)(~float_bits(z + 0.0f)) << 32) | (unsigned long long)(0xFFFFFFFEu - id);
}

struct Dims {
    int C, I, V, T, H, W;
};

__device__ __forceinline__ float4 load_clip(const float* clip, const Dims& d, int c, int i, int v)
{
    return reinterpret_cast<const float4*>(clip)[((size_t)c * d.I + i) * d.V + v];
}

// ---------------------------------------------------------------------------
// Raster: one thread per (camera, instance, triangle); snapped vertices,
// exact integer edge functions with a top-left rule, depth plane evaluated at
// pixel centres, 64-bit atomicMin of (depth, id).  Triangles with a vertex at
// w <= 0 are skipped (the cameras of this form see the mesh from outside).
__global__ __launch_bounds__(kBlock) void raster_tris_kernel(const float* __restrict__ clip,
                                                              const int* __restrict__ tris, Dims d,
                                                              unsigned long long* __restrict__ keys, int cull_back)
{
    const long long g = (long long)blockIdx.x * kBlock + threadIdx.x;
    const long long per_cam = (long long)d.I * d.T;
    if (g >= per_cam * d.C) return;
    const int c = (int)(g / per_cam);
    const int i = (int)((g % per_cam) / d.T);
    const int t = (int)(g % d.T);
    const float hw = (float)d.W * 0.5f, hh = (float)d.H * 0.5f;
    long long X[3], Y[3];
    float Z[3];
    for (int k = 0; k < 3; ++k) {
        const float4 p = load_clip(clip, d, c, i, tris[3 * t + k]);
        if (!(p.w > 0.0f)) return;
        const float xf = hw * (p.x / p.w) + hw, yf = hh * (p.y / p.w) + hh;
        if (!(fabsf(xf) < 1048576.0f && fabsf(yf) < 1048576.0f)) return;
        X[k] = (long long)rintf(xf * 256.0f);
        Y[k] = (long long)rintf(yf * 256.0f);
        Z[k] = p.z / p.w;
    }
    long long area = (X[1] - X[0]) * (Y[2] - Y[0]) - (Y[1] - Y[0]) * (X[2] - X[0]);
    if (area == 0) return;
    if (cull_back && area > 0) return;   // keep the winding the frame's back-face test keeps
    if (area < 0) {
        long long tx = X[1]; X[1] = X[2]; X[2] = tx;
        tx = Y[1]; Y[1] = Y[2]; Y[2] = tx;
        const float tz = Z[1]; Z[1] = Z[2]; Z[2] = tz;
        area = -area;
    }
    long long mnx = min(X[0], min(X[1], X[2])), mxx = max(X[0], max(X[1], X[2]));
    long long mny = min(Y[0], min(Y[1], Y[2])), mxy = max(Y[0], max(Y[1], Y[2]));
    const int px0 = (int)max((mnx - 128 + 255) >> 8, 0LL), px1 = (int)min((mxx - 128) >> 8, (long long)d.W - 1);
    const int py0 = (int)max((mny - 128 + 255) >> 8, 0LL), py1 = (int)min((mxy - 128) >> 8, (long long)d.H - 1);
    if (px0 > px1 || py0 > py1) return;
    long long ex[3], ey[3], ax[3], ay[3];
    bool tl[3];
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
    const uint32_t id = (uint32_t)(i * d.T + t + 1);
    unsigned long long* img = keys + (size_t)c * d.H * d.W;
    for (int py = py0; py <= py1; ++py) {
        const long long sy = (long long)py * 256 + 128;
        for (int px = px0; px <= px1; ++px) {
            const long long sx = (long long)px * 256 + 128;
            bool in = true;
            for (int k = 0; k < 3; ++k) {
                const long long e = ex[k] * (sy - ay[k]) - ey[k] * (sx - ax[k]);
                in = in && (e > 0 || (e == 0 && tl[k]));
            }
            if (!in) continue;
            const float z = Z[0] + (za * ((float)(sx - X[0]) * 0.00390625f) + zb * ((float)(sy - Y[0]) * 0.00390625f));
            if (!(z >= 0.0f && z <= 1.0f)) continue;
            (void)__hip_atomic_fetch_min(img + (size_t)py * d.W + px, make_key(z, id), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ __launch_bounds__(kBlock) void keys_to_ids_kernel(const unsigned long long* __restrict__ keys, size_t n,
                                                              int* __restrict__ ids, float* __restrict__ depth)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const unsigned long long k = keys[p];
    if (k == ~0ull) { ids[p] = 0; depth[p] = 0.0f; return; }
    ids[p] = (int)(0xFFFFFFFEu - (uint32_t)k);
    depth[p] = __uint_as_float(~(uint32_t)(k >> 32));
}

// ---------------------------------------------------------------------------
// Interpolation and antialiasing, templated on the type C of the clip
// positions (and of their gradient) and the type R of attributes, images,
// their gradients and the arithmetic.  (C, R) = (float, float), (double,
// double), or (double, float): the pixel-relative offsets that the results
// are ill-conditioned in (u below, the screen offsets of the antialiasing)
// are formed in C and then rounded to R.  The raster above always works on
// float clip positions.
template <typename T> struct V2 {
    T x, y;
};
template <typename T> struct V4 {
    T x, y, z, w;
};

template <typename T> __device__ __forceinline__ V4<T> load_clip4(const T* clip, const Dims& d, int c, int i, int v)
{
    const T* p = clip + (((size_t)c * d.I + i) * d.V + v) * 4;
    return V4<T>{ p[0], p[1], p[2], p[3] };
}

template <> __device__ __forceinline__ V4<float> load_clip4<float>(const float* clip, const Dims& d, int c, int i, int v)
{
    const float4 p = load_clip(clip, d, c, i, v);
    return V4<float>{ p.x, p.y, p.z, p.w };
}

template <typename T> __device__ __forceinline__ T cross2(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }

template <typename T> __device__ __forceinline__ T ndc_x(int px, const Dims& d)
{
    return (T(px) + T(0.5)) / (T(d.W) * T(0.5)) - T(1);
}
template <typename T> __device__ __forceinline__ T ndc_y(int py, const Dims& d)
{
    return (T(py) + T(0.5)) / (T(d.H) * T(0.5)) - T(1);
}

// Barycentrics at a pixel centre from the clip positions of a triangle.
template <typename R> struct Bary {
    V2<R> u[3];    // pixel-relative (x - xn w, y - yn w) per vertex
    R l[3], S;
    R lam[3];
    int v[3];
};

// An id outside 1..I*T or a vertex index outside 0..V-1 makes the pixel
// background: no read leaves the buffers whatever the caller's ids and tris.
__device__ __forceinline__ bool id_in_range(int id, const Dims& d) { return id > 0 && id <= d.I * d.T; }
__device__ __forceinline__ bool vertex_in_range(int v, const Dims& d) { return (unsigned)v < (unsigned)d.V; }

template <typename C, typename R>
__device__ __forceinline__ bool bary_at(const C* clip, const int* tris, const Dims& d, int c, int id, int px, int py,
                                        Bary<R>& b)
{
    if (!id_in_range(id, d)) return false;
    const int i = (id - 1) / d.T, t = (id - 1) % d.T;
    for (int k = 0; k < 3; ++k) {
        b.v[k] = tris[3 * t + k];
        if (!vertex_in_range(b.v[k], d)) return false;
    }
    const C xn = ndc_x<C>(px, d), yn = ndc_y<C>(py, d);
    for (int k = 0; k < 3; ++k) {
        const V4<C> p = load_clip4(clip, d, c, i, b.v[k]);
        b.u[k] = V2<R>{ R(fma(-xn, p.w, p.x)), R(fma(-yn, p.w, p.y)) };
    }
    for (int k = 0; k < 3; ++k) b.l[k] = cross2(b.u[(k + 1) % 3], b.u[(k + 2) % 3]);
    b.S = (b.l[0] + b.l[1]) + b.l[2];
    const R inv = R(1) / b.S;
    for (int k = 0; k < 3; ++k) b.lam[k] = b.l[k] * inv;
    return true;
}

template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void interp_fwd_kernel(const C* __restrict__ clip, const R* __restrict__ attr,
                                                             const int* __restrict__ tris, const int* __restrict__ ids,
                                                             Dims d, int A, R* __restrict__ out)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int id = ids[p];
    R* o = out + p * A;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    Bary<R> b;
    if (!bary_at(clip, tris, d, c, id, px, py, b)) {
        for (int a = 0; a < A; ++a) o[a] = R(0);
        return;
    }
    for (int a = 0; a < A; ++a)
        o[a] = (b.lam[0] * attr[b.v[0] * A + a] + b.lam[1] * attr[b.v[1] * A + a]) + b.lam[2] * attr[b.v[2] * A + a];
}

// d colour / d attr_k = lambda_k; d colour / d lambda_k = attr_k;
// d lambda_k / d l_j = (delta_kj - lambda_k) / S;
// l_k = u_{k+1} x u_{k+2}: dl_k/du_{k+1} = (u_{k+2}.y, -u_{k+2}.x), dl_k/du_{k+2} = (-u_{k+1}.y, u_{k+1}.x);
// u = (x - xn w, y - yn w): dL/dx = g.x, dL/dy = g.y, dL/dw = -(xn g.x + yn g.y).
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void interp_bwd_kernel(const C* __restrict__ clip, const R* __restrict__ attr,
                                                             const int* __restrict__ tris, const int* __restrict__ ids,
                                                             Dims d, int A, const R* __restrict__ g_out,
                                                             C* __restrict__ g_clip, R* __restrict__ g_attr)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int id = ids[p];
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    Bary<R> b;
    if (!bary_at(clip, tris, d, c, id, px, py, b)) return;
    const R xn = ndc_x<R>(px, d), yn = ndc_y<R>(py, d);
    const R* go = g_out + p * A;
    R glam[3] = { R(0), R(0), R(0) };
    for (int a = 0; a < A; ++a) {
        for (int k = 0; k < 3; ++k) {
            atomicAdd(&g_attr[b.v[k] * A + a], b.lam[k] * go[a]);
            glam[k] += go[a] * attr[b.v[k] * A + a];
        }
    }
    const R dot = (glam[0] * b.lam[0] + glam[1] * b.lam[1]) + glam[2] * b.lam[2];
    V2<R> gu[3] = { { R(0), R(0) }, { R(0), R(0) }, { R(0), R(0) } };
    const R inv = R(1) / b.S;
    for (int k = 0; k < 3; ++k) {
        const R gl = (glam[k] - dot) * inv;
        const V2<R> u1 = b.u[(k + 1) % 3], u2 = b.u[(k + 2) % 3];
        gu[(k + 1) % 3].x += gl * u2.y;
        gu[(k + 1) % 3].y -= gl * u2.x;
        gu[(k + 2) % 3].x -= gl * u1.y;
        gu[(k + 2) % 3].y += gl * u1.x;
    }
    const int i = (id - 1) / d.T;
    for (int k = 0; k < 3; ++k) {
        C* g = g_clip + (((size_t)c * d.I + i) * d.V + b.v[k]) * 4;
        atomicAdd(&g[0], C(gu[k].x));
        atomicAdd(&g[1], C(gu[k].y));
        atomicAdd(&g[3], C(-(xn * gu[k].x + yn * gu[k].y)));
    }
}

// ---------------------------------------------------------------------------
// Antialiasing.  A pair is (p, q) with q the right (dir 0) or lower (dir 1)
// neighbour.  pair_eval returns false when the pair does not blend; else the
// target pixel, alpha, and what the backward needs.  Screen positions are
// taken relative to the front pixel's centre (the back centre is then one
// unit step away), which keeps the edge functions exact to the rounding of
// the small offsets rather than of the screen size.
template <typename R> struct PairEval {
    size_t f, bpx;          // front and back pixel indices
    bool to_front;          // blend into the front pixel (t < 1/2) or the back one
    R t, alpha;
    int i, va, vb;          // instance and the crossed edge's vertices
    R ef, eb;               // edge functions at the two centres
    V2<R> sa, sb;           // screen offsets of the edge's vertices from the front centre
    V2<R> cb;               // back centre relative to the front one (a unit step)
    V4<R> pa, pb;           // clip positions of the edge's vertices
};

// hw (x / w - xc), hh (y / w - yc): screen offset in pixels from the centre
// with NDC (xc, yc), formed in C and rounded to R
template <typename C, typename R> __device__ __forceinline__ V2<R> screen_rel(V4<C> p, C xc, C yc, const Dims& d)
{
    const C hw = C(d.W) * C(0.5), hh = C(d.H) * C(0.5);
    return V2<R>{ R(hw * (fma(-xc, p.w, p.x) / p.w)), R(hh * (fma(-yc, p.w, p.y) / p.w)) };
}

template <typename R> __device__ __forceinline__ R edge_fn(V2<R> a, V2<R> b, V2<R> p)
{
    return (b.x - a.x) * (p.y - a.y) - (b.y - a.y) * (p.x - a.x);
}

template <typename C, typename R>
__device__ bool pair_eval(const C* clip, const int* tris, const int* ids, const float* depth, const Dims& d, int c,
                          int px, int py, int dir, PairEval<R>& e)
{
    const int qx = px + (dir == 0), qy = py + (dir == 1);
    if (qx >= d.W || qy >= d.H) return false;
    const size_t base = (size_t)c * d.H * d.W;
    const size_t ip = base + (size_t)py * d.W + px, iq = base + (size_t)qy * d.W + qx;
    const int idp = id_in_range(ids[ip], d) ? ids[ip] : 0, idq = id_in_range(ids[iq], d) ? ids[iq] : 0;
    if (idp == idq) return false;
    bool p_front;
    if (idp == 0) p_front = false;
    else if (idq == 0) p_front = true;
    else p_front = depth[ip] >= depth[iq];
    e.f = p_front ? ip : iq;
    e.bpx = p_front ? iq : ip;
    const int idf = p_front ? idp : idq;
    const int fx = p_front ? px : qx, fy = p_front ? py : qy;
    const R step = p_front ? R(1) : R(-1);
    e.cb = dir == 0 ? V2<R>{ step, R(0) } : V2<R>{ R(0), step };
    e.i = (idf - 1) / d.T;
    const int t = (idf - 1) % d.T;
    const C xc = ndc_x<C>(fx, d), yc = ndc_y<C>(fy, d);
    for (int k = 0; k < 3; ++k)
        if (!vertex_in_range(tris[3 * t + k], d)) return false;
    V4<C> P[3];
    V2<R> S[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = load_clip4(clip, d, c, e.i, tris[3 * t + k]);
        if (!(P[k].w > C(0))) return false;
        S[k] = screen_rel<C, R>(P[k], xc, yc, d);
    }
    // orientation: interior on the positive side of every edge (k, k+1)
    const R area = edge_fn(S[0], S[1], S[2]);
    if (area == R(0)) return false;
    const R sgn = area > R(0) ? R(1) : R(-1);
    const V2<R> zero{ R(0), R(0) };
    R best = R(2);
    int kb = -1;
    R bef = R(0), beb = R(0);
    for (int k = 0; k < 3; ++k) {
        const R ef = sgn * edge_fn(S[k], S[(k + 1) % 3], zero);
        const R eb = sgn * edge_fn(S[k], S[(k + 1) % 3], e.cb);
        if (!(ef >= R(0) && eb <= R(0) && ef > eb)) continue;
        const R tt = ef / (ef - eb);
        if (tt < best) { best = tt; kb = k; bef = ef; beb = eb; }
    }
    if (kb < 0) return false;
    e.t = best;
    e.to_front = best < R(0.5);
    e.alpha = e.to_front ? R(0.5) - best : best - R(0.5);
    e.va = tris[3 * t + kb];
    e.vb = tris[3 * t + (kb + 1) % 3];
    const V4<C> pa = P[kb], pb = P[(kb + 1) % 3];
    e.pa = V4<R>{ R(pa.x), R(pa.y), R(pa.z), R(pa.w) };
    e.pb = V4<R>{ R(pb.x), R(pb.y), R(pb.z), R(pb.w) };
    e.sa = S[kb];
    e.sb = S[(kb + 1) % 3];
    e.ef = sgn * bef;      // unsigned-orientation values, re-signed in the backward
    e.eb = sgn * beb;
    return true;
}

template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void aa_fwd_kernel(const R* __restrict__ color, const C* __restrict__ clip,
                                                         const int* __restrict__ tris, const int* __restrict__ ids,
                                                         const float* __restrict__ depth, Dims d, int A,
                                                         R* __restrict__ out)
{
    // gather: pixel p sums the blends of its four pairs that target it
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    R acc[8];
    for (int a = 0; a < A; ++a) acc[a] = color[p * A + a];
    const int sx[4] = { px, px - 1, px, px }, sy[4] = { py, py, py, py - 1 }, dr[4] = { 0, 0, 1, 1 };
    for (int q = 0; q < 4; ++q) {
        if (sx[q] < 0 || sy[q] < 0) continue;
        PairEval<R> e;
        if (!pair_eval(clip, tris, ids, depth, d, c, sx[q], sy[q], dr[q], e)) continue;
        const size_t tgt = e.to_front ? e.f : e.bpx, src = e.to_front ? e.bpx : e.f;
        if (tgt != p) continue;
        for (int a = 0; a < A; ++a) acc[a] += e.alpha * (color[src * A + a] - color[tgt * A + a]);
    }
    for (int a = 0; a < A; ++a) out[p * A + a] = acc[a];
}

template <typename C, typename R>
__device__ __forceinline__ void add_clip_grad(C* g_clip, const Dims& d, int c, int i, int v, R gx, R gy, R gw)
{
    C* g = g_clip + (((size_t)c * d.I + i) * d.V + v) * 4;
    atomicAdd(&g[0], C(gx));
    atomicAdd(&g[1], C(gy));
    atomicAdd(&g[3], C(gw));
}

// One thread per pair.  out_tgt = c_tgt + alpha (c_src - c_tgt):
//   d/dc_tgt = (1 - alpha), d/dc_src = alpha (the identity part of out_tgt is
//   propagated by the caller), d/dalpha = c_src - c_tgt;
//   alpha = +-(1/2 - t), t = ef / (ef - eb), e(P) = cross(sb - sa, P - sa).
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void aa_bwd_kernel(const R* __restrict__ color, const C* __restrict__ clip,
                                                         const int* __restrict__ tris, const int* __restrict__ ids,
                                                         const float* __restrict__ depth, Dims d, int A,
                                                         const R* __restrict__ g_out, R* __restrict__ g_color,
                                                         C* __restrict__ g_clip)
{
    const size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (g >= 2 * npx) return;
    const int dir = (int)(g / npx);
    const size_t p = g % npx;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    PairEval<R> e;
    if (!pair_eval(clip, tris, ids, depth, d, c, px, py, dir, e)) return;
    const size_t tgt = e.to_front ? e.f : e.bpx, src = e.to_front ? e.bpx : e.f;
    R galpha = R(0);
    for (int a = 0; a < A; ++a) {
        const R go = g_out[tgt * A + a];
        atomicAdd(&g_color[tgt * A + a], -e.alpha * go);
        atomicAdd(&g_color[src * A + a], e.alpha * go);
        galpha += go * (color[src * A + a] - color[tgt * A + a]);
    }
    const R gt = e.to_front ? -galpha : galpha;   // alpha = 1/2 - t or t - 1/2
    // t = ef / (ef - eb): dt/def = -eb / (ef - eb)^2, dt/deb = ef / (ef - eb)^2
    const R den = e.ef - e.eb;
    const R gef = gt * (-e.eb / (den * den)), geb = gt * (e.ef / (den * den));
    // e(P) = (bx - ax)(Py - ay) - (by - ay)(Px - ax):
    // de/dax = by - Py, de/day = Px - bx, de/dbx = Py - ay, de/dby = ax - Px
    const V2<R> A2 = e.sa, B2 = e.sb;
    R gax = R(0), gay = R(0), gbx = R(0), gby = R(0);
    const V2<R> Ps[2] = { { R(0), R(0) }, e.cb };
    const R ge[2] = { gef, geb };
    for (int k = 0; k < 2; ++k) {
        gax += ge[k] * (B2.y - Ps[k].y);
        gay += ge[k] * (Ps[k].x - B2.x);
        gbx += ge[k] * (Ps[k].y - A2.y);
        gby += ge[k] * (A2.x - Ps[k].x);
    }
    // screen offset s = (hw (x / w - xc), hh (y / w - yc)): ds/dx = hw / w, ds/dw = -hw x / w^2
    const R hw = R(d.W) * R(0.5), hh = R(d.H) * R(0.5);
    add_clip_grad<C, R>(g_clip, d, c, e.i, e.va, gax * hw / e.pa.w, gay * hh / e.pa.w,
                        -(gax * hw * e.pa.x + gay * hh * e.pa.y) / (e.pa.w * e.pa.w));
    add_clip_grad<C, R>(g_clip, d, c, e.i, e.vb, gbx * hw / e.pb.w, gby * hh / e.pb.w,
                        -(gbx * hw * e.pb.x + gby * hh * e.pb.y) / (e.pb.w * e.pb.w));
}

// ---------------------------------------------------------------------------
// Vertex transform: clip[c, i, v] = A[c, i] (p_v, 1) with A[c, i] = VP_c M_i
// (row-major 4x4).  One block covers kBlock consecutive vertices of one
// (camera, instance), so the matrix is uniform across the block.
template <typename C>
__global__ __launch_bounds__(kBlock) void transform_fwd_kernel(const C* __restrict__ pos, const C* __restrict__ A,
                                                                int V, C* __restrict__ clip)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    const int ci = blockIdx.y;
    if (v >= V) return;
    const C* m = A + (size_t)ci * 16;
    const C x = pos[3 * v], y = pos[3 * v + 1], z = pos[3 * v + 2];
    C* o = clip + ((size_t)ci * V + v) * 4;
    for (int r = 0; r < 4; ++r) o[r] = fma(m[4 * r], x, fma(m[4 * r + 1], y, fma(m[4 * r + 2], z, m[4 * r + 3])));
}

// g_pos[v] = sum over (c, i) of A[c, i][:, :3]^T g_clip[c, i, v]; one thread per vertex.
template <typename C>
__global__ __launch_bounds__(kBlock) void transform_bwd_pos_kernel(const C* __restrict__ g_clip, const C* __restrict__ A,
                                                                    int V, int CI, C* __restrict__ g_pos)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    C gx = C(0), gy = C(0), gz = C(0);
    for (int ci = 0; ci < CI; ++ci) {
        const C* g = g_clip + ((size_t)ci * V + v) * 4;
        const C* m = A + (size_t)ci * 16;
        const C g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
        gx += ((m[0] * g0 + m[4] * g1) + m[8] * g2) + m[12] * g3;
        gy += ((m[1] * g0 + m[5] * g1) + m[9] * g2) + m[13] * g3;
        gz += ((m[2] * g0 + m[6] * g1) + m[10] * g2) + m[14] * g3;
    }
    g_pos[3 * v] = gx;
    g_pos[3 * v + 1] = gy;
    g_pos[3 * v + 2] = gz;
}

// g_A[c, i] += sum over the block's vertices of g_clip[c, i, v] (p_v, 1)^T:
// per thread 16 products, summed across the wave with lane shuffles, then
// across the block's waves in LDS, then one atomic per entry and block.
template <typename C>
__global__ __launch_bounds__(kBlock) void transform_bwd_mat_kernel(const C* __restrict__ g_clip, const C* __restrict__ pos,
                                                                    int V, C* __restrict__ g_A)
{
    constexpr int kWaves = kBlock / 64;
    __shared__ C part[kWaves][16];
    const int v = blockIdx.x * kBlock + threadIdx.x;
    const int ci = blockIdx.y;
    C acc[16];
    if (v < V) {
        const C* g = g_clip + ((size_t)ci * V + v) * 4;
        const C ph[4] = { pos[3 * v], pos[3 * v + 1], pos[3 * v + 2], C(1) };
        for (int r = 0; r < 4; ++r) {
            const C gr = g[r];
            for (int k = 0; k < 4; ++k) acc[4 * r + k] = gr * ph[k];
        }
    } else {
        for (int k = 0; k < 16; ++k) acc[k] = C(0);
    }
    for (int k = 0; k < 16; ++k)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 16; ++k) part[wave][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 16) {
        C t = C(0);
        for (int w = 0; w < kWaves; ++w) t += part[w][threadIdx.x];
        atomicAdd(&g_A[(size_t)ci * 16 + threadIdx.x], t);
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void copy_kernel(const R* __restrict__ a, R* __restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

inline unsigned blocks(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

} // namespace

// ---------------------------------------------------------------------------
void raster_tris(const float* clip, const int* tris, int C, int I, int V, int T, int H, int W, int cull_back,
                 unsigned long long* keys, int* ids, float* depth, hipStream_t st)
{
    const Dims d{ C, I, V, T, H, W };
    const size_t npx = (size_t)C * H * W;
    const size_t ntri = (size_t)C * I * T;
    if (ntri && npx) raster_tris_kernel<<<blocks(ntri), kBlock, 0, st>>>(clip, tris, d, keys, cull_back);
    if (npx) keys_to_ids_kernel<<<blocks(npx), kBlock, 0, st>>>(keys, npx, ids, depth);
}

template <typename C, typename R>
static void interp_fwd_t(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T,
                         int H, int W, int A, R* out, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    if ((size_t)nc * H * W == 0) return;
    interp_fwd_kernel<C, R><<<blocks((size_t)nc * H * W), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, out);
}

template <typename C, typename R>
static void interp_bwd_t(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T,
                         int H, int W, int A, const R* g_out, C* g_clip, R* g_attr, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    if ((size_t)nc * H * W == 0) return;
    interp_bwd_kernel<C, R><<<blocks((size_t)nc * H * W), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, g_out, g_clip,
                                                                           g_attr);
}

template <typename C, typename R>
static void aa_fwd_t(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,
                     int V, int T, int H, int W, int A, R* out, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    if ((size_t)nc * H * W == 0) return;
    aa_fwd_kernel<C, R><<<blocks((size_t)nc * H * W), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, out);
}

template <typename C, typename R>
static void aa_bwd_t(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,
                     int V, int T, int H, int W, int A, const R* g_out, R* g_color, C* g_clip, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    const size_t n = (size_t)nc * H * W * A;
    if (n == 0) return;
    copy_kernel<R><<<blocks(n), kBlock, 0, st>>>(g_out, g_color, n);   // identity part of every output
    aa_bwd_kernel<C, R><<<blocks(2 * (size_t)nc * H * W), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, g_out,
                                                                           g_color, g_clip);
}

template <typename C>
static void transform_fwd_t(const C* pos, const C* A, int CI, int V, C* clip, hipStream_t st)
{
    if (CI == 0 || V == 0) return;
    transform_fwd_kernel<C><<<dim3(blocks((size_t)V), CI), kBlock, 0, st>>>(pos, A, V, clip);
}

template <typename C>
static void transform_bwd_t(const C* g_clip, const C* pos, const C* A, int CI, int V, C* g_pos, C* g_A, hipStream_t st)
{
    if (V == 0) return;
    transform_bwd_pos_kernel<C><<<blocks((size_t)V), kBlock, 0, st>>>(g_clip, A, V, CI, g_pos);
    if (CI) transform_bwd_mat_kernel<C><<<dim3(blocks((size_t)V), CI), kBlock, 0, st>>>(g_clip, pos, V, g_A);
}

void transform_fwd(const float* pos, const float* A, int CI, int V, float* clip, hipStream_t st)
{
    transform_fwd_t<float>(pos, A, CI, V, clip, st);
}
void transform_fwd(const double* pos, const double* A, int CI, int V, double* clip, hipStream_t st)
{
    transform_fwd_t<double>(pos, A, CI, V, clip, st);
}
void transform_bwd(const float* g_clip, const float* pos, const float* A, int CI, int V, float* g_pos, float* g_A,
                   hipStream_t st)
{
    transform_bwd_t<float>(g_clip, pos, A, CI, V, g_pos, g_A, st);
}
void transform_bwd(const double* g_clip, const double* pos, const double* A, int CI, int V, double* g_pos,
                   double* g_A, hipStream_t st)
{
    transform_bwd_t<double>(g_clip, pos, A, CI, V, g_pos, g_A, st);
}

#define VKGV_DIFF_LAUNCHERS(C, R)                                                                                     \
    void interp_fwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, R* out, hipStream_t st)                                                              \
    {                                                                                                                  \
        interp_fwd_t<C, R>(clip, attr, tris, ids, nc, I, V, T, H, W, A, out, st);                                      \
    }                                                                                                                  \
    void interp_bwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, const R* g_out, C* g_clip, R* g_attr, hipStream_t st)                                \
    {                                                                                                                  \
        interp_bwd_t<C, R>(clip, attr, tris, ids, nc, I, V, T, H, W, A, g_out, g_clip, g_attr, st);                    \
    }                                                                                                                  \
    void aa_fwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, R* out, hipStream_t st)                                             \
    {                                                                                                                  \
        aa_fwd_t<C, R>(color, clip, tris, ids, depth, nc, I, V, T, H, W, A, out, st);                                  \
    }                                                                                                                  \
    void aa_bwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, const R* g_out, R* g_color, C* g_clip, hipStream_t st)              \
    {                                                                                                                  \
        aa_bwd_t<C, R>(color, clip, tris, ids, depth, nc, I, V, T, H, W, A, g_out, g_color, g_clip, st);               \
    }

VKGV_DIFF_LAUNCHERS(float, float)
VKGV_DIFF_LAUNCHERS(double, double)
VKGV_DIFF_LAUNCHERS(double, float)
#undef VKGV_DIFF_LAUNCHERS

} // namespace vkgv_diff
