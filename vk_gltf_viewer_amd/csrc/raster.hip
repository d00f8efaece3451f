// Visibility-buffer rasterization: the mesh shader (visbuffer.mesh.glsl:30-104),
// the fixed-function raster state of the visbuffer pass
// (application.cpp:326-340, 772-861) and the ID write (visbuffer.frag.glsl:36).
//
// Raster rules (identical in every path below):
//  - vertices snapped to 1/256 pixel (round half to even), viewport [0,W]x[0,H];
//  - coverage at pixel centres with exact integer edge functions and a
//    top-left tie rule;
//  - depth = plane through the snapped vertices (set up in double, evaluated
//    in float), kept iff 0 <= z <= 1 (the near/far clip planes);
//  - triangles with a vertex at w <= 0 or beyond the 2^21-pixel guard band are
//    clipped in clip space (near plane, guard planes) and fanned;
//  - key = (~bits(z) << 32) | (0xFFFFFFFE - id), atomicMin: reversed-Z
//    GREATER_OR_EQUAL with ties to the later primitive.
//
// Work split: a persistent grid of 4-wave blocks; each wave takes one visible
// meshlet at a time from a cursor, transforms its <= 64 vertices into LDS,
// back-face culls its <= 124 triangles (2 per lane), then rasterizes the small
// ones cooperatively: a wave prefix sum over their bounding-box pixel counts
// spreads the pixels evenly over the 64 lanes.  Large and clipped triangles
// go to a queue rasterized by a tiled kernel (16x16 pixels per block).
#include "internal.hpp"

namespace vkgv {
namespace {

constexpr int kWaves = 4;
constexpr uint32_t kBlock = 64 * kWaves;
constexpr float kGuardPx = 2097152.0f;       // 2^21 pixels
constexpr float kClipGuardPx = 1048576.0f;   // guard clip planes at 2^20 pixels
constexpr unsigned long long kTileMask = (1ull << 40) - 1ull;

struct SVtx { int x, y; float z; uint32_t ok; };

struct TriSetup {
    int x[3], y[3];
    float z[3];
    int minx, maxx, miny, maxy;
    int px0, py0, px1, py1;
    long long area;
};

struct SmallTri {
    int C[3], A[3], B[3];
    float za, zb, z0, oxf, oyf;
    uint32_t org;   // px0 | py0 << 16
    uint32_t dim;   // bw | bh << 8
    uint32_t id;
};

struct RasterArgs {
    SceneDev s;
    Mat4 vp;
    const uint32_t* list;
    const uint32_t* list_count;
    uint32_t* cursor;
    unsigned long long* keys;
    uint32_t W, H;
    uint32_t flags;
    int small_lim;                  // small path iff bbox extent < small_lim (1/256 pixel units)
    BigTri* big;
    uint32_t* big_start;
    uint32_t big_cap;
    unsigned long long* big_ctr;
    uint32_t* clipped;
    uint32_t* overflow;
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool project_vtx(float4 c, float hw, float hh, SVtx& o)
{
    if (!(c.w > 0.0f)) return false;
    const float xd = c.x / c.w, yd = c.y / c.w, zd = c.z / c.w;
    const float xf = hw * xd + hw, yf = hh * yd + hh;
    if (!(fabsf(xf) < kGuardPx && fabsf(yf) < kGuardPx)) return false;
    o.x = (int)rintf(xf * 256.0f);
    o.y = (int)rintf(yf * 256.0f);
    o.z = zd;
    return true;
}

// Orientation (positive area), bounding box, pixel range clamped to the
// viewport.  Returns false for zero area or an empty pixel range.
__device__ __forceinline__ bool setup_tri(SVtx a, SVtx b, SVtx c, int W, int H, TriSetup& t)
{
    long long area = (long long)(b.x - a.x) * (long long)(c.y - a.y) - (long long)(b.y - a.y) * (long long)(c.x - a.x);
    if (area == 0) return false;
    if (area < 0) { const SVtx tmp = b; b = c; c = tmp; area = -area; }
    t.x[0] = a.x; t.x[1] = b.x; t.x[2] = c.x;
    t.y[0] = a.y; t.y[1] = b.y; t.y[2] = c.y;
    t.z[0] = a.z; t.z[1] = b.z; t.z[2] = c.z;
    t.area = area;
    t.minx = min(a.x, min(b.x, c.x)); t.maxx = max(a.x, max(b.x, c.x));
    t.miny = min(a.y, min(b.y, c.y)); t.maxy = max(a.y, max(b.y, c.y));
    t.px0 = max((t.minx - 128 + 255) >> 8, 0);
    t.px1 = min((t.maxx - 128) >> 8, W - 1);
    t.py0 = max((t.miny - 128 + 255) >> 8, 0);
    t.py1 = min((t.maxy - 128) >> 8, H - 1);
    return t.px0 <= t.px1 && t.py0 <= t.py1;
}

// Depth plane z(P) = z0 + za*dx + zb*dy through the snapped vertices.
__device__ __forceinline__ void depth_plane(const int* x, const int* y, const float* z, long long area, float& za, float& zb)
{
    const double e1x = (double)(x[1] - x[0]) / 256.0, e1y = (double)(y[1] - y[0]) / 256.0;
    const double e2x = (double)(x[2] - x[0]) / 256.0, e2y = (double)(y[2] - y[0]) / 256.0;
    const double dz1 = (double)z[1] - (double)z[0], dz2 = (double)z[2] - (double)z[0];
    const double det = (double)area / 65536.0;
    za = (float)((dz1 * e2y - dz2 * e1y) / det);
    zb = (float)((dz2 * e1x - dz1 * e2x) / det);
}

__device__ __forceinline__ void write_fragment(unsigned long long* keys, uint32_t W, int px, int py, float z, uint32_t id)
{
    const unsigned long long key = make_key(z, id);
    unsigned long long* p = keys + (size_t)py * W + px;
    if (key < *p) atomicMin(p, key);
}

__device__ void push_big(const TriSetup& t, uint32_t id, const RasterArgs& a)
{
    const uint32_t bw = (uint32_t)(t.px1 - t.px0 + 1), bh = (uint32_t)(t.py1 - t.py0 + 1);
    const uint32_t tiles = ((bw + 15u) >> 4) * ((bh + 15u) >> 4);
    const unsigned long long packed = atomicAdd(a.big_ctr, (1ull << 40) | (unsigned long long)tiles);
    const unsigned long long slot = packed >> 40;
    const unsigned long long start = packed & kTileMask;
    if (slot >= a.big_cap) { atomicOr(a.overflow, 1u); return; }
    BigTri b;
#pragma unroll
    for (int k = 0; k < 3; ++k) { b.x[k] = t.x[k]; b.y[k] = t.y[k]; b.z[k] = t.z[k]; }
    b.id = id;
    b.bb0 = (uint32_t)t.px0 | ((uint32_t)t.py0 << 16);
    b.bb1 = (uint32_t)t.px1 | ((uint32_t)t.py1 << 16);
    a.big[slot] = b;
    a.big_start[slot] = start > 0x7FFFFFFFull ? 0x7FFFFFFFu : (uint32_t)start;
    if (start + tiles > 0x7FFFFFFFull) atomicOr(a.overflow, 2u);
}

// Sutherland-Hodgman against a*x + b*y + c*z + d*w >= 0.
__device__ __forceinline__ int clip_plane(const float4* in, int n, float4* out, float pa, float pb, float pc, float pd)
{
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const float4 p = in[i];
        const float4 q = in[i + 1 < n ? i + 1 : 0];
        const float dp = ((pa * p.x + pb * p.y) + pc * p.z) + pd * p.w;
        const float dq = ((pa * q.x + pb * q.y) + pc * q.z) + pd * q.w;
        if (dp >= 0.0f) out[m++] = p;
        if ((dp >= 0.0f) != (dq >= 0.0f)) {
            const float t = dp / (dp - dq);
            float4 r;
            r.x = p.x + t * (q.x - p.x);
            r.y = p.y + t * (q.y - p.y);
            r.z = p.z + t * (q.z - p.z);
            r.w = p.w + t * (q.w - p.w);
            out[m++] = r;
        }
    }
    return m;
}

// Near-plane / guard-band clipping of one triangle; the fan goes to the
// tiled path.  Rare (triangles crossing w <= 0 or the 2^21-pixel guard band).
__device__ __noinline__ void clip_and_push(float4 c0, float4 c1, float4 c2, uint32_t id, const RasterArgs& a)
{
    float4 A[10], B[10];
    A[0] = c0; A[1] = c1; A[2] = c2;
    const float gx = kClipGuardPx / (float)a.W, gy = kClipGuardPx / (float)a.H;
    int n = 3;
    n = clip_plane(A, n, B, 0.0f, 0.0f, -1.0f, 1.0f);
    n = clip_plane(B, n, A, -1.0f, 0.0f, 0.0f, gx);
    n = clip_plane(A, n, B, 1.0f, 0.0f, 0.0f, gx);
    n = clip_plane(B, n, A, 0.0f, -1.0f, 0.0f, gy);
    n = clip_plane(A, n, B, 0.0f, 1.0f, 0.0f, gy);
    if (n < 3) return;
    atomicAdd(a.clipped, 1u);
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    SVtx sv[10];
    for (int i = 0; i < n; ++i)
        if (!project_vtx(B[i], hw, hh, sv[i])) return;
    for (int i = 1; i + 1 < n; ++i) {
        TriSetup t;
        if (setup_tri(sv[0], sv[i], sv[i + 1], (int)a.W, (int)a.H, t)) push_big(t, id, a);
    }
}

__device__ __forceinline__ uint32_t tri_byte(const uint32_t* words, uint32_t b)
{
    return (words[b >> 2] >> ((b & 3u) * 8u)) & 0xFFu;
}

__global__ __launch_bounds__(kBlock) void raster_kernel(RasterArgs a)
{
    __shared__ float4 s_clip[kWaves][64];
    __shared__ SVtx s_sv[kWaves][64];
    __shared__ uint32_t s_tri[kWaves][96];
    __shared__ SmallTri s_st[kWaves][128];
    __shared__ uint32_t s_pre[kWaves][128];

    const int w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t count = __builtin_amdgcn_readfirstlane(*a.list_count);
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    const int W = (int)a.W, H = (int)a.H;

    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(a.cursor, 1u);
        idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
        if (idx >= count) break;
        const uint32_t d = __builtin_amdgcn_readfirstlane(a.list[idx]);
        const vkgv_meshlet_draw dr = a.s.draws[d];
        const vkgv_primitive& prim = a.s.prims[dr.primitive_index];
        const vkgv_meshlet ml = reinterpret_cast<const vkgv_meshlet*>(prim.meshlet_buffer)[dr.meshlet_index];
        const vkgv_material& mat = a.s.materials[prim.material_index];
        const Mat4 m = load_mat4(a.s.transforms + 16ull * dr.transform_index);
        const Mat4 mvp = mat4_mul(a.vp, m);
        const float tdet = det4(m);
        const bool cull_faces = !mat.double_sided && !(a.flags & VKGV_NO_BACKFACE);
        const uint32_t vc = ml.vertex_count, tc = ml.triangle_count;

        // vertices: clip = (VP * M) * (pos, 1) (visbuffer.mesh.glsl:44-65)
        if (lane < vc) {
            const uint32_t vi = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer)[ml.vertex_offset + lane];
            const float* pp = reinterpret_cast<const float*>(prim.vertex_buffer) + (size_t)vi * 6;
            const float4 c = mat4_point(mvp, pp[0], pp[1], pp[2]);
            s_clip[w][lane] = c;
            SVtx sv;
            sv.x = 0; sv.y = 0; sv.z = 0.0f;
            sv.ok = project_vtx(c, hw, hh, sv) ? 1u : 0u;
            s_sv[w][lane] = sv;
        }
        {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(prim.primitive_index_buffer + ml.triangle_offset);
            const uint32_t ndw = (tc * 3u + 3u) >> 2;
            for (uint32_t k = lane; k < ndw; k += 64u) s_tri[w][k] = src[k];
        }
        wave_sync();

        uint32_t npix[2] = { 0u, 0u };
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            const uint32_t t = lane + 64u * (uint32_t)sl;
            if (t >= tc) continue;
            const uint32_t i0 = tri_byte(s_tri[w], 3u * t), i1 = tri_byte(s_tri[w], 3u * t + 1u),
                           i2 = tri_byte(s_tri[w], 3u * t + 2u);
            const float4 c0 = s_clip[w][i0], c1 = s_clip[w][i1], c2 = s_clip[w][i2];
            if (cull_faces) {
                const float det = det3_xyw(c0, c1, c2);
                const bool culled = tdet < 0.0f ? det < 0.0f : det > 0.0f;
                if (culled) continue;
            }
            const uint32_t id = (d << VKGV_TRIANGLE_BITS) | t;
            const SVtx v0 = s_sv[w][i0], v1 = s_sv[w][i1], v2 = s_sv[w][i2];
            if (!(v0.ok & v1.ok & v2.ok)) {
                clip_and_push(c0, c1, c2, id, a);
                continue;
            }
            TriSetup ts;
            if (!setup_tri(v0, v1, v2, W, H, ts)) continue;
            if (ts.maxx - ts.minx < a.small_lim && ts.maxy - ts.miny < a.small_lim) {
                SmallTri st;
                const int sx0 = ts.px0 * 256 + 128, sy0 = ts.py0 * 256 + 128;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int ka = (k + 1) % 3, kb = (k + 2) % 3;
                    const int ex = ts.x[kb] - ts.x[ka], ey = ts.y[kb] - ts.y[ka];
                    const bool tl = ey < 0 || (ey == 0 && ex > 0);
                    st.C[k] = ex * (sy0 - ts.y[ka]) - ey * (sx0 - ts.x[ka]) - (tl ? 0 : 1);
                    st.A[k] = -ey * 256;
                    st.B[k] = ex * 256;
                }
                depth_plane(ts.x, ts.y, ts.z, ts.area, st.za, st.zb);
                st.z0 = ts.z[0];
                st.oxf = (float)(sx0 - ts.x[0]) * 0.00390625f;
                st.oyf = (float)(sy0 - ts.y[0]) * 0.00390625f;
                st.org = (uint32_t)ts.px0 | ((uint32_t)ts.py0 << 16);
                const uint32_t bw = (uint32_t)(ts.px1 - ts.px0 + 1), bh = (uint32_t)(ts.py1 - ts.py0 + 1);
                st.dim = bw | (bh << 8);
                st.id = id;
                s_st[w][t] = st;
                npix[sl] = bw * bh;
            } else {
                push_big(ts, id, a);
            }
        }

        // exclusive prefix of pixel counts over the 128 triangle slots
        uint32_t x0 = npix[0], x1 = npix[1];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y0 = __shfl_up(x0, off), y1 = __shfl_up(x1, off);
            if (lane >= (uint32_t)off) { x0 += y0; x1 += y1; }
        }
        const uint32_t tot0 = __shfl(x0, 63), tot1 = __shfl(x1, 63);
        s_pre[w][lane] = x0 - npix[0];
        s_pre[w][64 + lane] = tot0 + x1 - npix[1];
        const uint32_t total = tot0 + tot1;
        wave_sync();

        for (uint32_t k = lane; k < total; k += 64u) {
            uint32_t lo = 0, hi = 127;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (s_pre[w][mid] <= k) lo = mid; else hi = mid - 1u;
            }
            const SmallTri& st = s_st[w][lo];
            const uint32_t r = k - s_pre[w][lo];
            const uint32_t bw = st.dim & 0xFFu;
            const uint32_t j = (uint32_t)(((float)r + 0.5f) * __builtin_amdgcn_rcpf((float)bw));
            const uint32_t i = r - j * bw;
            if (i >= bw) continue;   // never taken: (r + 0.5) / bw is at least 0.5/bw from an integer
            const int ii = (int)i, jj = (int)j;
            const int e0 = st.C[0] + st.A[0] * ii + st.B[0] * jj;
            const int e1 = st.C[1] + st.A[1] * ii + st.B[1] * jj;
            const int e2 = st.C[2] + st.A[2] * ii + st.B[2] * jj;
            if ((e0 | e1 | e2) < 0) continue;
            const float z = st.z0 + (st.za * (st.oxf + (float)ii) + st.zb * (st.oyf + (float)jj));
            if (!(z >= 0.0f && z <= 1.0f)) continue;
            write_fragment(a.keys, a.W, (int)(st.org & 0xFFFFu) + ii, (int)(st.org >> 16) + jj, z, st.id);
        }
        wave_sync();
    }
}

struct BigArgs {
    const BigTri* big;
    const uint32_t* start;
    uint32_t cap;
    unsigned long long* big_ctr;
    Counters* counters;
    unsigned long long* keys;
    uint32_t W;
};

__device__ __forceinline__ uint32_t big_tiles(const BigTri& t)
{
    const uint32_t bw = (t.bb1 & 0xFFFFu) - (t.bb0 & 0xFFFFu) + 1u, bh = (t.bb1 >> 16) - (t.bb0 >> 16) + 1u;
    return ((bw + 15u) >> 4) * ((bh + 15u) >> 4);
}

__global__ __launch_bounds__(256) void raster_big_kernel(BigArgs a)
{
    const unsigned long long packed = *a.big_ctr;
    const uint32_t raw = (uint32_t)(packed >> 40);
    const uint32_t count = raw < a.cap ? raw : a.cap;
    if (count == 0) return;
    unsigned long long tot = packed & kTileMask;
    if (raw > a.cap) tot = (unsigned long long)a.start[count - 1] + big_tiles(a.big[count - 1]);
    if (tot > 0x7FFFFFFFull) tot = 0x7FFFFFFFull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&a.counters->big_total_tris, count);
        atomicAdd(&a.counters->big_total_tiles, (uint32_t)tot);
    }
    const uint32_t total = (uint32_t)tot;
    for (uint32_t g = blockIdx.x; g < total; g += gridDim.x) {
        uint32_t lo = 0, hi = count - 1u;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (a.start[mid] <= g) lo = mid; else hi = mid - 1u;
        }
        const BigTri t = a.big[lo];
        const uint32_t local = g - a.start[lo];
        const int px0 = (int)(t.bb0 & 0xFFFFu), py0 = (int)(t.bb0 >> 16);
        const int px1 = (int)(t.bb1 & 0xFFFFu), py1 = (int)(t.bb1 >> 16);
        const uint32_t ntx = ((uint32_t)(px1 - px0 + 1) + 15u) >> 4;
        const uint32_t ty = local / ntx, tx = local - ty * ntx;
        const int px = px0 + (int)tx * 16 + (int)(threadIdx.x & 15u);
        const int py = py0 + (int)ty * 16 + (int)(threadIdx.x >> 4);
        if (px > px1 || py > py1) continue;
        const long long sx = (long long)px * 256 + 128, sy = (long long)py * 256 + 128;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ka = (k + 1) % 3, kb = (k + 2) % 3;
            const long long ex = (long long)t.x[kb] - t.x[ka], ey = (long long)t.y[kb] - t.y[ka];
            const bool tl = ey < 0 || (ey == 0 && ex > 0);
            const long long e = ex * (sy - t.y[ka]) - ey * (sx - t.x[ka]);
            inside = inside && (e > 0 || (e == 0 && tl));
        }
        if (!inside) continue;
        const long long area = ((long long)t.x[1] - t.x[0]) * ((long long)t.y[2] - t.y[0])
                               - ((long long)t.y[1] - t.y[0]) * ((long long)t.x[2] - t.x[0]);
        float za, zb;
        depth_plane(t.x, t.y, t.z, area, za, zb);
        const float dxf = (float)(sx - t.x[0]) * 0.00390625f;
        const float dyf = (float)(sy - t.y[0]) * 0.00390625f;
        const float z = t.z[0] + (za * dxf + zb * dyf);
        if (!(z >= 0.0f && z <= 1.0f)) continue;
        write_fragment(a.keys, a.W, px, py, z, t.id);
    }
}

} // namespace

void launch_raster(const SceneDev& s, const ViewDev& v, const Mat4& vp, uint32_t flags, uint32_t list,
                   uint32_t grid_blocks, uint32_t small_max_extent, hipStream_t st)
{
    RasterArgs a;
    a.s = s;
    a.vp = vp;
    a.list = v.visible[list];
    a.list_count = &v.counters->vis_count[list];
    a.cursor = &v.counters->raster_cursor[list];
    a.keys = v.keys;
    a.W = v.width;
    a.H = v.height;
    a.flags = flags;
    a.small_lim = (int)(small_max_extent * 256u);
    a.big = v.big;
    a.big_start = v.big_tile_start;
    a.big_cap = v.big_capacity;
    a.big_ctr = &v.counters->big;
    a.clipped = &v.counters->clipped;
    a.overflow = &v.counters->overflow;
    raster_kernel<<<grid_blocks, kBlock, 0, st>>>(a);
}

void launch_raster_big(const ViewDev& v, uint32_t grid_blocks, hipStream_t st)
{
    BigArgs a;
    a.big = v.big;
    a.start = v.big_tile_start;
    a.cap = v.big_capacity;
    a.big_ctr = &v.counters->big;
    a.counters = v.counters;
    a.keys = v.keys;
    a.W = v.width;
    raster_big_kernel<<<grid_blocks, 256, 0, st>>>(a);
}

} // namespace vkgv
