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
__device__ __forceinline__ void bary_from(const V4<C> P[3], C xn, C yn, Bary<R>& b)
{
    for (int k = 0; k < 3; ++k) b.u[k] = V2<R>{ R(fma(-xn, P[k].w, P[k].x)), R(fma(-yn, P[k].w, P[k].y)) };
    for (int k = 0; k < 3; ++k) b.l[k] = cross2(b.u[(k + 1) % 3], b.u[(k + 2) % 3]);
    b.S = (b.l[0] + b.l[1]) + b.l[2];
    const R inv = R(1) / b.S;
    for (int k = 0; k < 3; ++k) b.lam[k] = b.l[k] * inv;
}

// The triangle of an in-range id: its instance, vertices (false when one is
// out of range) and clip positions.
template <typename C>
__device__ __forceinline__ bool tri_of(const C* clip, const int* tris, const Dims& d, int c, int id, int& i, int& t,
                                       int vi[3], V4<C> P[3])
{
    i = (id - 1) / d.T;
    t = (id - 1) % d.T;
    for (int k = 0; k < 3; ++k) {
        vi[k] = tris[3 * t + k];
        if (!vertex_in_range(vi[k], d)) return false;
    }
    for (int k = 0; k < 3; ++k) P[k] = load_clip4(clip, d, c, i, vi[k]);
    return true;
}

template <typename C, typename R>
__device__ __forceinline__ bool bary_at(const C* clip, const int* tris, const Dims& d, int c, int id, int px, int py,
                                        Bary<R>& b)
{
    if (!id_in_range(id, d)) return false;
    int i, t;
    V4<C> P[3];
    if (!tri_of(clip, tris, d, c, id, i, t, b.v, P)) return false;
    bary_from(P, ndc_x<C>(px, d), ndc_y<C>(py, d), b);
    return true;
}

// ---------------------------------------------------------------------------
// Gathers.  The backward of interpolation and antialiasing runs in two
// passes.  A pixel pass computes, with every lane busy, each pixel's terms
// for the three corners of its triangle (plain stores, contrib).  A vertex
// pass, one thread per (camera, instance, vertex), walks the vertex's
// incident triangles (a CSR adjacency of the mesh) and sums the terms of the
// pixels in each triangle's screen box that carry the triangle's id.  No
// atomics, and the sums run in a fixed order: the gradients are the same
// bits on every run.
//
// The box holds every pixel whose centre lies in the triangle, with a margin
// of 1/64 px over the raster's 1/256-px snapping, so rasterize's ids always
// fall inside it.  A pixel outside its own triangle's box (ids from
// elsewhere), or of a triangle that is not gathered (a vertex at w <= 0 or not
// finite, a repeated vertex, or a box over kGatherMaxBox pixels), is a stray.
// The forwards flag each block of pixels that holds a stray; the backwards
// scatter those pixels' terms with atomics after the vertex pass, in a pass
// that skips unflagged blocks.
constexpr int kGatherMaxBox = 64;
constexpr int kGatherBatch = 8;   // incident triangles whose boxes are loaded together
constexpr int kPixBatch = 4;      // box pixels whose ids are loaded together

struct Box {
    int x0, x1, y0, y1;
};

// boxes[(c * I + i) * T + t]: x0 | x1 << 16 | y0 << 32 | y1 << 48, or kNoBox
// for a triangle that is not gathered (or has a vertex out of range); an
// empty box (x0 > x1) for one on no pixel.
constexpr unsigned long long kNoBox = ~0ull;

__device__ __forceinline__ unsigned long long pack_box(const Box& b)
{
    return (unsigned long long)(uint16_t)b.x0 | ((unsigned long long)(uint16_t)b.x1 << 16)
           | ((unsigned long long)(uint16_t)b.y0 << 32) | ((unsigned long long)(uint16_t)b.y1 << 48);
}

__device__ __forceinline__ Box unpack_box(unsigned long long v)
{
    return Box{ (int)(v & 0xFFFFu), (int)((v >> 16) & 0xFFFFu), (int)((v >> 32) & 0xFFFFu), (int)(v >> 48) };
}

// Index of (camera c, the instance and triangle of in-range id) in boxes and vis.
__device__ __forceinline__ size_t tri_index(const Dims& d, int c, int id)
{
    return (size_t)c * d.I * d.T + (size_t)(id - 1);
}

__device__ __forceinline__ bool finite_coord(float x) { return fabsf(x) < 1e30f; }

// The box, in float whatever C (the raster snaps float clip positions): the
// hardware reciprocal's error (1 ulp) is far inside the margin.
template <typename C>
__device__ __forceinline__ bool gather_box(const V4<C> P[3], const int vi[3], const Dims& d, Box& b)
{
    if (vi[0] == vi[1] || vi[1] == vi[2] || vi[0] == vi[2]) return false;
    const float hw = (float)d.W * 0.5f, hh = (float)d.H * 0.5f;
    float mnx = 0.0f, mxx = 0.0f, mny = 0.0f, mxy = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float w = (float)P[k].w;
        if (!(w > 0.0f)) return false;
        // pixel coordinates: the centre of pixel (px, py) is at (px, py)
        const float r = __builtin_amdgcn_rcpf(w);
        const float sx = fmaf((float)P[k].x * r, hw, hw - 0.5f), sy = fmaf((float)P[k].y * r, hh, hh - 0.5f);
        if (!(finite_coord(sx) && finite_coord(sy))) return false;
        mnx = k ? fminf(mnx, sx) : sx;
        mxx = k ? fmaxf(mxx, sx) : sx;
        mny = k ? fminf(mny, sy) : sy;
        mxy = k ? fmaxf(mxy, sy) : sy;
    }
    const float margin = 1.0f / 64.0f;
    const float fx0 = fmaxf(ceilf(mnx - margin), 0.0f), fx1 = fminf(floorf(mxx + margin), (float)(d.W - 1));
    const float fy0 = fmaxf(ceilf(mny - margin), 0.0f), fy1 = fminf(floorf(mxy + margin), (float)(d.H - 1));
    if (fx0 > fx1 || fy0 > fy1) {   // on no pixel
        b = Box{ 1, 0, 1, 0 };
        return true;
    }
    if ((fx1 - fx0 + 1.0f) * (fy1 - fy0 + 1.0f) > (float)kGatherMaxBox) return false;
    b = Box{ (int)fx0, (int)fx1, (int)fy0, (int)fy1 };
    return true;
}

// One thread per (camera, instance, triangle).
// Also clears vis, the per-triangle flags that the forward's pixel pass sets
// for the triangles with a gathered pixel.
template <typename C>
__global__ __launch_bounds__(kBlock) void tri_boxes_kernel(const C* __restrict__ clip, const int* __restrict__ tris,
                                                            Dims d, unsigned long long* __restrict__ boxes,
                                                            uint8_t* __restrict__ vis)
{
    const size_t g = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t n = (size_t)d.C * d.I * d.T;
    if (g >= n) return;
    const int t = (int)(g % d.T);
    const size_t ci = g / d.T;
    const int c = (int)(ci / d.I), i = (int)(ci % d.I);
    int ii, tt, vi[3];
    V4<C> P[3];
    Box b;
    const bool ok = tri_of(clip, tris, d, c, i * d.T + t + 1, ii, tt, vi, P) && gather_box(P, vi, d, b);
    boxes[g] = ok ? pack_box(b) : kNoBox;
    vis[g] = 0;
}

// Whether pixel (px, py) of camera c with id `id` is a stray (false for
// background).
__device__ __forceinline__ bool pixel_stray(const unsigned long long* boxes, const Dims& d, int c, int id, int px,
                                            int py)
{
    if (!id_in_range(id, d)) return false;
    const unsigned long long v = boxes[tri_index(d, c, id)];
    if (v == kNoBox) return true;
    const Box b = unpack_box(v);
    return !(px >= b.x0 && px <= b.x1 && py >= b.y0 && py <= b.y1);
}

// One flag per block of kBlock pixels: whether the block holds a stray.
// Every thread of the block must call it.
__device__ __forceinline__ void flag_block(bool stray, unsigned* __restrict__ flags)
{
    const int any = __syncthreads_or(stray ? 1 : 0);
    if (threadIdx.x == 0) flags[blockIdx.x] = any ? 1u : 0u;
}

// (camera-instance pair, vertex) of a vertex-pass thread; false past the end.
__device__ __forceinline__ bool gather_thread(const Dims& d, int& ci, int& v)
{
    const unsigned nvb = (unsigned)((d.V + (int)kBlock - 1) / (int)kBlock);
    ci = (int)(blockIdx.x / nvb);
    v = (int)((blockIdx.x % nvb) * kBlock + threadIdx.x);
    return v < d.V;
}

template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void interp_fwd_kernel(const C* __restrict__ clip, const R* __restrict__ attr,
                                                             const int* __restrict__ tris, const int* __restrict__ ids,
                                                             Dims d, int A, R* __restrict__ out,
                                                             const unsigned long long* __restrict__ boxes,
                                                             unsigned* __restrict__ stray_blocks,
                                                             uint8_t* __restrict__ vis)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    bool stray = false;
    if (p < npx) {
        const int id = ids[p];
        R* o = out + p * A;
        const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
        Bary<R> b;
        if (!bary_at(clip, tris, d, c, id, px, py, b)) {
            for (int a = 0; a < A; ++a) o[a] = R(0);
        } else {
            for (int a = 0; a < A; ++a)
                o[a] = (b.lam[0] * attr[b.v[0] * A + a] + b.lam[1] * attr[b.v[1] * A + a])
                       + b.lam[2] * attr[b.v[2] * A + a];
            stray = pixel_stray(boxes, d, c, id, px, py);
            if (!stray) vis[tri_index(d, c, id)] = 1;   // every writer stores the same value
        }
    }
    flag_block(stray, stray_blocks);
}

// d colour / d attr_k = lambda_k; d colour / d lambda_k = attr_k;
// d lambda_k / d l_j = (delta_kj - lambda_k) / S;
// l_k = u_{k+1} x u_{k+2}: dl_k/du_{k+1} = (u_{k+2}.y, -u_{k+2}.x), dl_k/du_{k+2} = (-u_{k+1}.y, u_{k+1}.x);
// u = (x - xn w, y - yn w): dL/dx = g.x, dL/dy = g.y, dL/dw = -(xn g.x + yn g.y).
// glam: dL/dlambda_k; gl: dL/dl_k = (glam_k - sum_j glam_j lambda_j) / S.
template <typename R> __device__ __forceinline__ void lambda_grad(const Bary<R>& b, const R glam[3], R gl[3])
{
    const R dot = (glam[0] * b.lam[0] + glam[1] * b.lam[1]) + glam[2] * b.lam[2];
    const R inv = R(1) / b.S;
    for (int k = 0; k < 3; ++k) gl[k] = (glam[k] - dot) * inv;
}

// dL/du_k from dL/dl: l_{k-1} = u_k x u_{k+1} and l_{k+1} = u_{k-1} x u_k.
template <typename R> __device__ __forceinline__ V2<R> corner_grad(const Bary<R>& b, const R gl[3], int k)
{
    const int kp = (k + 1) % 3, km = (k + 2) % 3;
    return V2<R>{ gl[km] * b.u[kp].y - gl[kp] * b.u[km].y, gl[kp] * b.u[km].x - gl[km] * b.u[kp].x };
}

// The interpolation terms of pixel p (false for a background pixel): per
// corner k, (dL/dx, dL/dy, dL/dw) of its clip position and lambda_k g_out.
template <typename C, typename R> struct InterpTerms {
    Bary<R> b;
    V2<R> gu[3];
    R gw[3];
};

template <typename C, typename R>
__device__ __forceinline__ bool interp_terms(const C* clip, const R* attr, const int* tris, const Dims& d, int A,
                                             int c, int id, int px, int py, const R* go, InterpTerms<C, R>& o)
{
    if (!bary_at(clip, tris, d, c, id, px, py, o.b)) return false;
    const R xn = ndc_x<R>(px, d), yn = ndc_y<R>(py, d);
    R glam[3] = { R(0), R(0), R(0) };
    for (int a = 0; a < A; ++a)
        for (int k = 0; k < 3; ++k) glam[k] += go[a] * attr[o.b.v[k] * A + a];
    R gl[3];
    lambda_grad(o.b, glam, gl);
    for (int k = 0; k < 3; ++k) {
        o.gu[k] = corner_grad(o.b, gl, k);
        o.gw[k] = -(xn * o.gu[k].x + yn * o.gu[k].y);
    }
    return true;
}

// Pixel pass: contrib[p][k] = (dL/dx, dL/dy, dL/dw, lambda_k g_out[p][0..A-1]),
// with the antialiasing pass's terms contrib_a[p][k] added to the first three
// when given (one load per pixel in the vertex pass of the combined backward).
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void interp_bwd_pixel_kernel(const C* __restrict__ clip,
                                                                   const R* __restrict__ attr,
                                                                   const int* __restrict__ tris,
                                                                   const int* __restrict__ ids, Dims d, int A,
                                                                   const R* __restrict__ g_out,
                                                                   const R* __restrict__ contrib_a,
                                                                   R* __restrict__ contrib)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    const R* go = g_out + p * A;
    InterpTerms<C, R> it;
    if (!interp_terms(clip, attr, tris, d, A, c, ids[p], px, py, go, it)) return;
    R ca[9];
    for (int k = 0; k < 9; ++k) ca[k] = contrib_a ? contrib_a[p * 9 + k] : R(0);
    R* o = contrib + p * 3 * (3 + A);
    for (int k = 0; k < 3; ++k) {
        R* ok = o + k * (3 + A);
        ok[0] = it.gu[k].x + ca[3 * k];
        ok[1] = it.gu[k].y + ca[3 * k + 1];
        ok[2] = it.gw[k] + ca[3 * k + 2];
        for (int a = 0; a < A; ++a) ok[3 + a] = it.b.lam[k] * go[a];
    }
}

// The stray pass: a stray pixel's whole interpolation gradient, with atomics.
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void interp_bwd_stray_kernel(const C* __restrict__ clip,
                                                                   const R* __restrict__ attr,
                                                                   const int* __restrict__ tris,
                                                                   const int* __restrict__ ids, Dims d, int A,
                                                                   const R* __restrict__ g_out,
                                                                   const unsigned long long* __restrict__ boxes,
                                                                   const unsigned* __restrict__ stray_blocks,
                                                                   C* __restrict__ g_clip, R* __restrict__ g_attr)
{
    if (stray_blocks[blockIdx.x] == 0u) return;   // uniform over the block
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int id = ids[p];
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    if (!pixel_stray(boxes, d, c, id, px, py)) return;
    const R* go = g_out + p * A;
    InterpTerms<C, R> it;
    if (!interp_terms(clip, attr, tris, d, A, c, id, px, py, go, it)) return;
    const int i = (id - 1) / d.T;
    for (int k = 0; k < 3; ++k) {
        for (int a = 0; a < A; ++a) atomicAdd(&g_attr[it.b.v[k] * A + a], it.b.lam[k] * go[a]);
        C* g = g_clip + (((size_t)c * d.I + i) * d.V + it.b.v[k]) * 4;
        atomicAdd(&g[0], C(it.gu[k].x));
        atomicAdd(&g[1], C(it.gu[k].y));
        atomicAdd(&g[3], C(it.gw[k]));
    }
}

// The vertex pass: g_clip[c, i, v] (written, z = 0) and, with the
// interpolation terms, this camera-instance pair's attribute gradient
// part[c, i, v, A] (reduced over the pairs by attr_reduce_kernel).  The
// interpolation terms are contrib_i[p][k][3 + A], the antialiasing ones
// contrib_a[p][k][3]; MODE says which are summed (1 or 2).  Triangles
// without a gathered pixel (vis 0) are skipped.
template <typename C, typename R, int MODE>
__global__ __launch_bounds__(kBlock) void vertex_gather_kernel(const int* __restrict__ ids, Dims d, int A,
                                                                const int* __restrict__ adj_off,
                                                                const int* __restrict__ adj,
                                                                const unsigned long long* __restrict__ boxes,
                                                                const uint8_t* __restrict__ vis,
                                                                const R* __restrict__ contrib_i,
                                                                const R* __restrict__ contrib_a,
                                                                C* __restrict__ g_clip, R* __restrict__ part)
{
    int ci, v;
    if (!gather_thread(d, ci, v)) return;
    const int i = ci % d.I;
    const size_t base = (size_t)(ci / d.I) * d.H * d.W;
    const unsigned long long* cbox = boxes + (size_t)ci * d.T;
    const uint8_t* cvis = vis + (size_t)ci * d.T;
    C gx = C(0), gy = C(0), gw = C(0);
    R ga[8];
    for (int a = 0; a < A; ++a) ga[a] = R(0);
    const int j0 = adj_off[v], j1 = adj_off[v + 1];
    for (int jb = j0; jb < j1; jb += kGatherBatch) {
        // the batch's corners, then its boxes and flags: independent loads in flight together
        int corner[kGatherBatch];
        unsigned long long bx[kGatherBatch];
        uint8_t seen[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) corner[u] = jb + u < j1 ? adj[jb + u] : -1;
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            bx[u] = corner[u] >= 0 ? cbox[corner[u] / 3] : kNoBox;
            seen[u] = corner[u] >= 0 ? cvis[corner[u] / 3] : 0;
        }
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
            if (!seen[u] || bx[u] == kNoBox) continue;   // no gathered pixel: nothing to sum
            const int t = corner[u] / 3, k = corner[u] % 3, id = i * d.T + t + 1;
            const Box b = unpack_box(bx[u]);
            const int bw = b.x1 - b.x0 + 1, n = bw * (b.y1 - b.y0 + 1);
            // the box's ids kPixBatch at a time, all loads issued before the first compare
            for (int q0 = 0; q0 < n; q0 += kPixBatch) {
                size_t ip[kPixBatch];
                int idq[kPixBatch];
                int qx = b.x0 + q0 % bw, qy = b.y0 + q0 / bw;
#pragma unroll
                for (int w = 0; w < kPixBatch; ++w) {
                    ip[w] = base + (size_t)qy * d.W + qx;
                    idq[w] = q0 + w < n ? ids[ip[w]] : 0;
                    if (++qx > b.x1) {
                        qx = b.x0;
                        ++qy;
                    }
                }
#pragma unroll
                for (int w = 0; w < kPixBatch; ++w) {
                    if (idq[w] != id) continue;
                    if (MODE & 1) {
                        const R* o = contrib_i + (ip[w] * 3 + k) * (3 + A);
                        gx += C(o[0]);
                        gy += C(o[1]);
                        gw += C(o[2]);
                        for (int a = 0; a < A; ++a) ga[a] += o[3 + a];
                    }
                    if (MODE & 2) {
                        const R* o = contrib_a + ip[w] * 9 + 3 * k;
                        gx += C(o[0]);
                        gy += C(o[1]);
                        gw += C(o[2]);
                    }
                }
            }
        }
    }
    C* g = g_clip + ((size_t)ci * d.V + v) * 4;
    g[0] = gx;
    g[1] = gy;
    g[2] = C(0);
    g[3] = gw;
    if (MODE & 1) {
        R* o = part + ((size_t)ci * d.V + v) * A;
        for (int a = 0; a < A; ++a) o[a] = ga[a];
    }
}

// g_attr[v, a] = sum over the camera-instance pairs of part[ci, v, a], in pair order.
template <typename R>
__global__ __launch_bounds__(kBlock) void attr_reduce_kernel(const R* __restrict__ part, int CI, size_t VA,
                                                              R* __restrict__ g_attr)
{
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= VA) return;
    R s = R(0);
    for (int ci = 0; ci < CI; ++ci) s += part[(size_t)ci * VA + j];
    g_attr[j] = s;
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
    int ka;                 // corner of the front triangle at va (vb is the next one)
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

// The blend of a pair whose front pixel (centre at NDC (xc, yc)) holds the
// triangle with vertices vi and clip positions P of instance i, the back
// centre at cb from the front one: the first edge of the triangle that the
// segment from the front centre to the back one leaves through.  Fills every
// field but f and bpx.
template <typename C, typename R>
__device__ __forceinline__ bool front_eval(const V4<C> P[3], const int vi[3], int i, C xc, C yc, V2<R> cb,
                                           const Dims& d, PairEval<R>& e)
{
    V2<R> S[3];
    for (int k = 0; k < 3; ++k) {
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
        const R eb = sgn * edge_fn(S[k], S[(k + 1) % 3], cb);
        if (!(ef >= R(0) && eb <= R(0) && ef > eb)) continue;
        const R tt = ef / (ef - eb);
        if (tt < best) { best = tt; kb = k; bef = ef; beb = eb; }
    }
    if (kb < 0) return false;
    e.cb = cb;
    e.i = i;
    e.t = best;
    e.to_front = best < R(0.5);
    e.alpha = e.to_front ? R(0.5) - best : best - R(0.5);
    e.ka = kb;
    e.va = vi[kb];
    e.vb = vi[(kb + 1) % 3];
    const V4<C> pa = P[kb], pb = P[(kb + 1) % 3];
    e.pa = V4<R>{ R(pa.x), R(pa.y), R(pa.z), R(pa.w) };
    e.pb = V4<R>{ R(pb.x), R(pb.y), R(pb.z), R(pb.w) };
    e.sa = S[kb];
    e.sb = S[(kb + 1) % 3];
    e.ef = sgn * bef;      // unsigned-orientation values, re-signed in the backward
    e.eb = sgn * beb;
    return true;
}

// The front pixel of pair (a, b), b the right (dir 0) or lower (dir 1)
// neighbour of a, given their ids (0 when out of range) with ida != idb:
// the one with an id, else the one with the larger depth (a on ties).
__device__ __forceinline__ bool a_is_front(int ida, int idb, float da, float db)
{
    if (ida == 0) return false;
    if (idb == 0) return true;
    return da >= db;
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
    const bool p_front = a_is_front(idp, idq, depth[ip], depth[iq]);
    e.f = p_front ? ip : iq;
    e.bpx = p_front ? iq : ip;
    const int idf = p_front ? idp : idq;
    const int fx = p_front ? px : qx, fy = p_front ? py : qy;
    const R step = p_front ? R(1) : R(-1);
    const V2<R> cb = dir == 0 ? V2<R>{ step, R(0) } : V2<R>{ R(0), step };
    const int i = (idf - 1) / d.T, t = (idf - 1) % d.T;
    int vi[3];
    for (int k = 0; k < 3; ++k) {
        vi[k] = tris[3 * t + k];
        if (!vertex_in_range(vi[k], d)) return false;
    }
    V4<C> P[3];
    for (int k = 0; k < 3; ++k) P[k] = load_clip4(clip, d, c, i, vi[k]);
    return front_eval<C, R>(P, vi, i, ndc_x<C>(fx, d), ndc_y<C>(fy, d), cb, d, e);
}

// The four pairs of pixel (px, py), as (a, dir) with b the right (dir 0) or
// lower (dir 1) neighbour of a: (p, right), (left, p), (p, down), (up, p).
__device__ __forceinline__ bool pair_of(int q, int px, int py, int& ax, int& ay, int& dir)
{
    ax = q == 1 ? px - 1 : px;
    ay = q == 3 ? py - 1 : py;
    dir = q >> 1;
    return ax >= 0 && ay >= 0;
}

// stray_blocks may be null (the flags and vis come from elsewhere).
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void aa_fwd_kernel(const R* __restrict__ color, const C* __restrict__ clip,
                                                         const int* __restrict__ tris, const int* __restrict__ ids,
                                                         const float* __restrict__ depth, Dims d, int A,
                                                         R* __restrict__ out,
                                                         const unsigned long long* __restrict__ boxes,
                                                         unsigned* __restrict__ stray_blocks, uint8_t* __restrict__ vis)
{
    // gather: pixel p sums the blends of its four pairs that target it
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    bool stray = false;
    if (p < npx) {
        const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
        R acc[8];
        for (int a = 0; a < A; ++a) acc[a] = color[p * A + a];
        for (int q = 0; q < 4; ++q) {
            int ax, ay, dir;
            if (!pair_of(q, px, py, ax, ay, dir)) continue;
            PairEval<R> e;
            if (!pair_eval(clip, tris, ids, depth, d, c, ax, ay, dir, e)) continue;
            const size_t tgt = e.to_front ? e.f : e.bpx, src = e.to_front ? e.bpx : e.f;
            if (tgt != p) continue;
            for (int a = 0; a < A; ++a) acc[a] += e.alpha * (color[src * A + a] - color[tgt * A + a]);
        }
        for (int a = 0; a < A; ++a) out[p * A + a] = acc[a];
        if (stray_blocks) {
            const int id = ids[p];
            stray = pixel_stray(boxes, d, c, id, px, py);
            if (id_in_range(id, d) && !stray) vis[tri_index(d, c, id)] = 1;
        }
    }
    if (stray_blocks) flag_block(stray, stray_blocks);   // uniform over the grid
}

// One blend's backward.  out_tgt = c_tgt + alpha (c_src - c_tgt):
//   d/dc_tgt = (1 - alpha), d/dc_src = alpha, d/dalpha = c_src - c_tgt;
//   alpha = +-(1/2 - t), t = ef / (ef - eb), e(P) = cross(sb - sa, P - sa).
// galpha = sum over channels of g_out[tgt] (c_src - c_tgt); returns
// dL/d(x, y, w) of the crossed edge's vertices a = va and b = vb.
template <typename R> struct EdgeGrad {
    R ax, ay, aw, bx, by, bw;
};

template <typename R> __device__ __forceinline__ EdgeGrad<R> edge_grad(const PairEval<R>& e, R galpha, const Dims& d)
{
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
    EdgeGrad<R> g;
    g.ax = gax * hw / e.pa.w;
    g.ay = gay * hh / e.pa.w;
    g.aw = -(gax * hw * e.pa.x + gay * hh * e.pa.y) / (e.pa.w * e.pa.w);
    g.bx = gbx * hw / e.pb.w;
    g.by = gby * hh / e.pb.w;
    g.bw = -(gbx * hw * e.pb.x + gby * hh * e.pb.y) / (e.pb.w * e.pb.w);
    return g;
}

template <typename R>
__device__ __forceinline__ R blend_galpha(const PairEval<R>& e, const R* color, const R* g_out, int A)
{
    const size_t tgt = e.to_front ? e.f : e.bpx, src = e.to_front ? e.bpx : e.f;
    R galpha = R(0);
    for (int a = 0; a < A; ++a) galpha += g_out[tgt * A + a] * (color[src * A + a] - color[tgt * A + a]);
    return galpha;
}

template <typename C, typename R>
__device__ __forceinline__ void add_clip_grad(C* g_clip, const Dims& d, int c, int i, int v, R gx, R gy, R gw)
{
    C* g = g_clip + (((size_t)c * d.I + i) * d.V + v) * 4;
    atomicAdd(&g[0], C(gx));
    atomicAdd(&g[1], C(gy));
    atomicAdd(&g[3], C(gw));
}

// Pixel pass.  g_color (written): g_out (the identity part) plus the pixel's
// share of the blends of its four pairs, -alpha g_out[p] where it is the
// target and alpha g_out[tgt] where it is the source.  contrib[p][k] (for
// every pixel with an id): dL/d(x, y, w) of corner k of the pixel's triangle
// from the blends whose front pixel it is, summed over its pairs.
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void aa_bwd_pixel_kernel(const R* __restrict__ color, const C* __restrict__ clip,
                                                               const int* __restrict__ tris,
                                                               const int* __restrict__ ids,
                                                               const float* __restrict__ depth, Dims d, int A,
                                                               const R* __restrict__ g_out, R* __restrict__ g_color,
                                                               R* __restrict__ contrib)
{
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    R acc[8];
    for (int a = 0; a < A; ++a) acc[a] = g_out[p * A + a];
    R cg[9];
    for (int k = 0; k < 9; ++k) cg[k] = R(0);
    for (int q = 0; q < 4; ++q) {
        int ax, ay, dir;
        if (!pair_of(q, px, py, ax, ay, dir)) continue;
        PairEval<R> e;
        if (!pair_eval(clip, tris, ids, depth, d, c, ax, ay, dir, e)) continue;
        const size_t tgt = e.to_front ? e.f : e.bpx;
        if (tgt == p)
            for (int a = 0; a < A; ++a) acc[a] += -e.alpha * g_out[tgt * A + a];
        else
            for (int a = 0; a < A; ++a) acc[a] += e.alpha * g_out[tgt * A + a];
        if (e.f != p) continue;
        const EdgeGrad<R> g = edge_grad(e, blend_galpha(e, color, g_out, A), d);
        const int ka = 3 * e.ka, kb = 3 * ((e.ka + 1) % 3);
        cg[ka] += g.ax;
        cg[ka + 1] += g.ay;
        cg[ka + 2] += g.aw;
        cg[kb] += g.bx;
        cg[kb + 1] += g.by;
        cg[kb + 2] += g.bw;
    }
    for (int a = 0; a < A; ++a) g_color[p * A + a] = acc[a];
    if (id_in_range(ids[p], d))
        for (int k = 0; k < 9; ++k) contrib[p * 9 + k] = cg[k];
}

// The stray pass: the edge terms of the blends whose front pixel is a stray,
// scattered with atomics.
template <typename C, typename R>
__global__ __launch_bounds__(kBlock) void aa_bwd_stray_kernel(const R* __restrict__ color, const C* __restrict__ clip,
                                                               const int* __restrict__ tris,
                                                               const int* __restrict__ ids,
                                                               const float* __restrict__ depth, Dims d, int A,
                                                               const R* __restrict__ g_out,
                                                               const unsigned long long* __restrict__ boxes,
                                                               const unsigned* __restrict__ stray_blocks,
                                                               C* __restrict__ g_clip)
{
    if (stray_blocks[blockIdx.x] == 0u) return;   // uniform over the block
    const size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x;
    const size_t npx = (size_t)d.C * d.H * d.W;
    if (p >= npx) return;
    const int c = (int)(p / ((size_t)d.H * d.W)), py = (int)((p / d.W) % d.H), px = (int)(p % d.W);
    if (!pixel_stray(boxes, d, c, ids[p], px, py)) return;
    for (int q = 0; q < 4; ++q) {
        int ax, ay, dir;
        if (!pair_of(q, px, py, ax, ay, dir)) continue;
        PairEval<R> e;
        if (!pair_eval(clip, tris, ids, depth, d, c, ax, ay, dir, e) || e.f != p) continue;
        const EdgeGrad<R> g = edge_grad(e, blend_galpha(e, color, g_out, A), d);
        add_clip_grad<C, R>(g_clip, d, c, e.i, e.va, g.ax, g.ay, g.aw);
        add_clip_grad<C, R>(g_clip, d, c, e.i, e.vb, g.bx, g.by, g.bw);
    }
}

// ---------------------------------------------------------------------------
// Vertex adjacency of a mesh (CSR): off[v] .. off[v + 1] - 1 index the
// corners 3 t + k with tris[3 t + k] = v, in increasing order.  Vertex
// indices out of range are left out (their triangles are background).
__global__ __launch_bounds__(kBlock) void adj_zero_kernel(int* __restrict__ a, size_t n)
{
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j < n) a[j] = 0;
}

__global__ __launch_bounds__(kBlock) void adj_count_kernel(const int* __restrict__ tris, size_t n, int V,
                                                            int* __restrict__ cnt)
{
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const int v = tris[j];
    if ((unsigned)v < (unsigned)V) atomicAdd(&cnt[v], 1);
}

// One block: off[1..V] holds the counts, replaced by their inclusive sums.
constexpr int kAdjScanThreads = 1024, kAdjScanPer = 4;
__global__ __launch_bounds__(kAdjScanThreads) void adj_scan_kernel(int* __restrict__ off, int V)
{
    __shared__ int s_w[kAdjScanThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int carry = 0;
    for (int base = 0; base < V; base += kAdjScanThreads * kAdjScanPer) {
        int x[kAdjScanPer], sum = 0;
        for (int u = 0; u < kAdjScanPer; ++u) {
            const int idx = base + tid * kAdjScanPer + u;
            x[u] = idx < V ? off[1 + idx] : 0;
            sum += x[u];
        }
        int incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) s_w[w] = incl;
        __syncthreads();
        int wbase = 0, total = 0;
        for (int q = 0; q < kAdjScanThreads / 64; ++q) {
            if (q < w) wbase += s_w[q];
            total += s_w[q];
        }
        int run = carry + wbase + incl - sum;
        for (int u = 0; u < kAdjScanPer; ++u) {
            const int idx = base + tid * kAdjScanPer + u;
            run += x[u];
            if (idx < V) off[1 + idx] = run;
        }
        carry += total;
        __syncthreads();   // s_w reused by the next round
    }
}

__global__ __launch_bounds__(kBlock) void adj_fill_kernel(const int* __restrict__ tris, size_t n, int V,
                                                           const int* __restrict__ off, int* __restrict__ cursor,
                                                           int* __restrict__ adj)
{
    const size_t j = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const int v = tris[j];
    if ((unsigned)v < (unsigned)V) adj[off[v] + atomicAdd(&cursor[v], 1)] = (int)j;
}

// Insertion sort of each vertex's corners (its valence, at most a few dozen
// in a sane mesh): the gathers then sum in a fixed order.
__global__ __launch_bounds__(kBlock) void adj_sort_kernel(const int* __restrict__ off, int V, int* __restrict__ adj)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= V) return;
    const int b = off[v], e = off[v + 1];
    for (int j = b + 1; j < e; ++j) {
        const int x = adj[j];
        int k = j - 1;
        while (k >= b && adj[k] > x) {
            adj[k + 1] = adj[k];
            --k;
        }
        adj[k + 1] = x;
    }
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

size_t stray_flag_count(int nc, int H, int W) { return blocks((size_t)nc * H * W); }

void mesh_adjacency(const int* tris, int T, int V, int* off, int* adj, hipStream_t st)
{
    if (V <= 0) return;
    const size_t n = (size_t)3 * T;
    int* cursor = off + V + 1;
    adj_zero_kernel<<<blocks((size_t)2 * V + 1), kBlock, 0, st>>>(off, (size_t)2 * V + 1);
    if (n) adj_count_kernel<<<blocks(n), kBlock, 0, st>>>(tris, n, V, off + 1);
    adj_scan_kernel<<<1, kAdjScanThreads, 0, st>>>(off, V);
    if (n) adj_fill_kernel<<<blocks(n), kBlock, 0, st>>>(tris, n, V, off, cursor, adj);
    adj_sort_kernel<<<blocks((size_t)V), kBlock, 0, st>>>(off, V, adj);
}

size_t tri_boxes_words(int nc, int I, int T)
{
    const size_t n = (size_t)nc * I * T;
    return n + (n + 7) / 8;
}

// The flag bytes that follow the box words.
static uint8_t* vis_of(unsigned long long* boxes, int nc, int I, int T)
{
    return reinterpret_cast<uint8_t*>(boxes + (size_t)nc * I * T);
}
static const uint8_t* vis_of(const unsigned long long* boxes, int nc, int I, int T)
{
    return reinterpret_cast<const uint8_t*>(boxes + (size_t)nc * I * T);
}

template <typename C>
static void tri_boxes_t(const C* clip, const int* tris, int nc, int I, int V, int T, int H, int W,
                        unsigned long long* boxes, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    const size_t n = (size_t)nc * I * T;
    if (n) tri_boxes_kernel<C><<<blocks(n), kBlock, 0, st>>>(clip, tris, d, boxes, vis_of(boxes, nc, I, T));
}

void tri_boxes(const float* clip, const int* tris, int nc, int I, int V, int T, int H, int W,
               unsigned long long* boxes, hipStream_t st)
{
    tri_boxes_t<float>(clip, tris, nc, I, V, T, H, W, boxes, st);
}
void tri_boxes(const double* clip, const int* tris, int nc, int I, int V, int T, int H, int W,
               unsigned long long* boxes, hipStream_t st)
{
    tri_boxes_t<double>(clip, tris, nc, I, V, T, H, W, boxes, st);
}

// Blocks of a vertex pass over (camera-instance pair, vertex).
static unsigned gather_blocks(int nc, int I, int V) { return blocks((size_t)V) * (unsigned)(nc * I); }

template <typename C, typename R>
static void interp_fwd_t(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T,
                         int H, int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks,
                         hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    if ((size_t)nc * H * W == 0) return;
    interp_fwd_kernel<C, R><<<blocks((size_t)nc * H * W), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, out, boxes,
                                                                           stray_blocks, vis_of(boxes, nc, I, T));
}

template <typename C, typename R>
static void interp_bwd_t(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T,
                         int H, int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks,
                         const int* adj_off, const int* adj, const R* g_out, R* contrib, R* part, C* g_clip, R* g_attr,
                         hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    const size_t npx = (size_t)nc * H * W;
    if (V == 0) return;
    if (npx)
        interp_bwd_pixel_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, g_out, nullptr,
                                                                      contrib);
    if (nc * I)
        vertex_gather_kernel<C, R, 1><<<gather_blocks(nc, I, V), kBlock, 0, st>>>(
            ids, d, A, adj_off, adj, boxes, vis_of(boxes, nc, I, T), contrib, nullptr, g_clip, part);
    attr_reduce_kernel<R><<<blocks((size_t)V * A), kBlock, 0, st>>>(part, nc * I, (size_t)V * A, g_attr);
    if (npx)
        interp_bwd_stray_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, g_out, boxes,
                                                                      stray_blocks, g_clip, g_attr);
}

template <typename C, typename R>
static void aa_fwd_t(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,
                     int V, int T, int H, int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks,
                     hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    if ((size_t)nc * H * W == 0) return;
    aa_fwd_kernel<C, R><<<blocks((size_t)nc * H * W), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, out, boxes,
                                                                       stray_blocks, vis_of(boxes, nc, I, T));
}

template <typename C, typename R>
static void aa_bwd_t(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,
                     int V, int T, int H, int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks,
                     const int* adj_off, const int* adj, const R* g_out, R* contrib, R* g_color, C* g_clip,
                     hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    const size_t npx = (size_t)nc * H * W;
    if (npx)
        aa_bwd_pixel_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, g_out, g_color,
                                                                  contrib);
    if (V && nc * I)
        vertex_gather_kernel<C, R, 2><<<gather_blocks(nc, I, V), kBlock, 0, st>>>(
            ids, d, A, adj_off, adj, boxes, vis_of(boxes, nc, I, T), nullptr, contrib, g_clip, nullptr);
    if (npx && V)
        aa_bwd_stray_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, g_out, boxes,
                                                                  stray_blocks, g_clip);
}

// antialias(interpolate(...)) as one backward: the antialiasing pixel pass
// gives the colour gradient and its edge terms, the interpolation pixel pass
// takes both (adding the edge terms to its own), and one vertex pass sums
// them into a single g_clip.
template <typename C, typename R>
static void shade_bwd_t(const R* color, const C* clip, const R* attr, const int* tris, const int* ids,
                        const float* depth, int nc, int I, int V, int T, int H, int W, int A,
                        const unsigned long long* boxes, const unsigned* stray_blocks, const int* adj_off,
                        const int* adj, const R* g_out, R* g_color, R* contrib_a, R* contrib_i, R* part, C* g_clip,
                        R* g_attr, hipStream_t st)
{
    const Dims d{ nc, I, V, T, H, W };
    const size_t npx = (size_t)nc * H * W;
    if (V == 0) return;
    if (npx) {
        aa_bwd_pixel_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, g_out, g_color,
                                                                  contrib_a);
        interp_bwd_pixel_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, g_color, contrib_a,
                                                                      contrib_i);
    }
    if (nc * I)
        vertex_gather_kernel<C, R, 1><<<gather_blocks(nc, I, V), kBlock, 0, st>>>(
            ids, d, A, adj_off, adj, boxes, vis_of(boxes, nc, I, T), contrib_i, nullptr, g_clip, part);
    attr_reduce_kernel<R><<<blocks((size_t)V * A), kBlock, 0, st>>>(part, nc * I, (size_t)V * A, g_attr);
    if (npx) {
        interp_bwd_stray_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(clip, attr, tris, ids, d, A, g_color, boxes,
                                                                      stray_blocks, g_clip, g_attr);
        aa_bwd_stray_kernel<C, R><<<blocks(npx), kBlock, 0, st>>>(color, clip, tris, ids, depth, d, A, g_out, boxes,
                                                                  stray_blocks, g_clip);
    }
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
                    int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks, hipStream_t st)           \
    {                                                                                                                  \
        interp_fwd_t<C, R>(clip, attr, tris, ids, nc, I, V, T, H, W, A, out, boxes, stray_blocks, st);                 \
    }                                                                                                                  \
    void interp_bwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks, const int* adj_off,   \
                    const int* adj, const R* g_out, R* contrib, R* part, C* g_clip, R* g_attr, hipStream_t st)         \
    {                                                                                                                  \
        interp_bwd_t<C, R>(clip, attr, tris, ids, nc, I, V, T, H, W, A, boxes, stray_blocks, adj_off, adj, g_out,      \
                           contrib, part, g_clip, g_attr, st);                                                         \
    }                                                                                                                  \
    void aa_fwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks,          \
                hipStream_t st)                                                                                        \
    {                                                                                                                  \
        aa_fwd_t<C, R>(color, clip, tris, ids, depth, nc, I, V, T, H, W, A, out, boxes, stray_blocks, st);             \
    }                                                                                                                  \
    void aa_bwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks,      \
                const int* adj_off, const int* adj, const R* g_out, R* contrib, R* g_color, C* g_clip,                 \
                hipStream_t st)                                                                                        \
    {                                                                                                                  \
        aa_bwd_t<C, R>(color, clip, tris, ids, depth, nc, I, V, T, H, W, A, boxes, stray_blocks, adj_off, adj, g_out,  \
                       contrib, g_color, g_clip, st);                                                                  \
    }                                                                                                                  \
    void shade_bwd(const R* color, const C* clip, const R* attr, const int* tris, const int* ids, const float* depth,  \
                   int nc, int I, int V, int T, int H, int W, int A, const unsigned long long* boxes,                  \
                   const unsigned* stray_blocks, const int* adj_off, const int* adj, const R* g_out, R* g_color,       \
                   R* contrib_a, R* contrib_i, R* part, C* g_clip, R* g_attr, hipStream_t st)                          \
    {                                                                                                                  \
        shade_bwd_t<C, R>(color, clip, attr, tris, ids, depth, nc, I, V, T, H, W, A, boxes, stray_blocks, adj_off,     \
                          adj, g_out, g_color, contrib_a, contrib_i, part, g_clip, g_attr, st);                        \
    }

VKGV_DIFF_LAUNCHERS(float, float)
VKGV_DIFF_LAUNCHERS(double, double)
VKGV_DIFF_LAUNCHERS(double, float)
#undef VKGV_DIFF_LAUNCHERS

} // namespace vkgv_diff
