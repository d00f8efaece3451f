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
// Work split.
//  raster_kernel  persistent grid of 4-wave blocks; each wave takes one visible
//                 meshlet at a time (the first by its wave index, the rest from
//                 an atomic cursor), transforms its <= 64 vertices into LDS,
//                 back-face culls its <= 124 triangles (2 per lane) and
//                 rasterizes the small ones cooperatively: a wave prefix sum
//                 over their bounding-box pixel counts spreads the pixels evenly
//                 over the 64 lanes.  Large triangles go to the big queue (one
//                 64-bit atomic reserves the slot and the 64-tile chunks of the
//                 triangle's 32x32-pixel bounding-box tiles), triangles needing
//                 clipping to the clip queue.
//  clip_kernel    clips the queued triangles; their fans join the big queue.
//  bin_kernel     one wave per chunk: stores the triangle's setup (chunk 0),
//                 tests 64 tiles, writes a (tile, triangle) pair per tile that
//                 can hold a covered pixel centre and counts it in its tile; the
//                 last block to finish scans the per-tile counts.
//  bin_fill       scatters the pairs into per-tile triangle lists.
//  tile_kernel    kTileSplit blocks per screen tile, each rasterizing its share
//                 of the tile's triangles into a 32x32 key tile in LDS (LDS
//                 atomicMin), then merging it into the visibility buffer with
//                 one global atomicMin per covered pixel.
#include "internal.hpp"

namespace vkgv {
namespace {

constexpr int kWaves = 4;
constexpr uint32_t kBlock = 64 * kWaves;
constexpr float kGuardPx = 2097152.0f;       // 2^21 pixels
constexpr float kClipGuardPx = 1048576.0f;   // guard clip planes at 2^20 pixels

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
    uint32_t* big_chunk;
    uint32_t big_cap;
    ClipTri* clipq;
    uint32_t clip_cap;
    Counters* ctr;
    const uint32_t* slot_base;      // first resolve-cache slot of this list (null: 0)
    TriCache* tri_cache;
    float4* slot_albedo;
    uint32_t* draw_slot;
    uint32_t cache_slots;
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

// Global 64-bit minimum without a returned value (the wave does not wait).
__device__ __forceinline__ void key_min(unsigned long long* p, unsigned long long key)
{
    (void)__hip_atomic_fetch_min(p, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void write_fragment(unsigned long long* keys, uint32_t W, int px, int py, float z, uint32_t id)
{
    const unsigned long long key = make_key(z, id);
    unsigned long long* p = keys + (size_t)py * W + px;
    if (key < *p) key_min(p, key);
}

// Appends a big triangle: one 64-bit atomic gives its slot (high word) and the
// first of the 64-tile chunks its bounding box needs (low word).
__device__ __forceinline__ void push_big(const TriSetup& t, uint32_t id, BigTri* big, uint32_t* big_chunk, uint32_t cap,
                                         Counters* ctr)
{
    const uint32_t ntx = (uint32_t)(t.px1 / (int)kBinTile - t.px0 / (int)kBinTile + 1);
    const uint32_t nty = (uint32_t)(t.py1 / (int)kBinTile - t.py0 / (int)kBinTile + 1);
    const uint32_t chunks = (ntx * nty + 63u) >> 6;
    const unsigned long long packed = atomicAdd(&ctr->big_packed, (1ull << 32) | (unsigned long long)chunks);
    const uint32_t slot = (uint32_t)(packed >> 32);
    if (slot >= cap) { atomicOr(&ctr->overflow, 1u); return; }
    big_chunk[slot] = (uint32_t)packed;
    BigTri b;
#pragma unroll
    for (int k = 0; k < 3; ++k) { b.x[k] = t.x[k]; b.y[k] = t.y[k]; b.z[k] = t.z[k]; }
    b.id = id;
    b.bb0 = (uint32_t)t.px0 | ((uint32_t)t.py0 << 16);
    b.bb1 = (uint32_t)t.px1 | ((uint32_t)t.py1 << 16);
    big[slot] = b;
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
    __shared__ uint32_t s_col[kWaves][64];
    const uint32_t count = __builtin_amdgcn_readfirstlane(*a.list_count);
    const uint32_t slot0 = a.slot_base ? __builtin_amdgcn_readfirstlane(*a.slot_base) : 0u;
    const uint32_t nwaves = gridDim.x * kWaves;
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    const int W = (int)a.W, H = (int)a.H;

    uint32_t idx = blockIdx.x * kWaves + (uint32_t)w;   // first meshlet: static
    while (idx < count) {
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
        // resolve-cache slot of this meshlet (none when the cache is full)
        const uint32_t slot = slot0 + idx;
        const bool cached = slot < a.cache_slots;
        if (lane == 0) {
            a.draw_slot[d] = cached ? slot : 0xFFFFFFFFu;
            if (cached)
                a.slot_albedo[slot] = make_float4(mat.albedo_factor[0], mat.albedo_factor[1], mat.albedo_factor[2],
                                                  mat.albedo_factor[3]);
        }

        // vertices: clip = (VP * M) * (pos, 1) (visbuffer.mesh.glsl:44-65)
        if (lane < vc) {
            const uint32_t vi = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer)[ml.vertex_offset + lane];
            const float* pp = reinterpret_cast<const float*>(prim.vertex_buffer) + (size_t)vi * 6;
            const float4 c = mat4_point(mvp, pp[0], pp[1], pp[2]);
            s_clip[w][lane] = c;
            s_col[w][lane] = reinterpret_cast<const uint32_t*>(pp)[3];
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
            if (cached) {
                // r[k] = v_{k+1} x v_{k+2} over the clip (x, y, w) of the vertices
                TriCache tcache;
                const float4 cv[3] = { c0, c1, c2 };
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float4 p = cv[(k + 1) % 3], q = cv[(k + 2) % 3];
                    tcache.r[k][0] = p.y * q.w - p.w * q.y;
                    tcache.r[k][1] = p.w * q.x - p.x * q.w;
                    tcache.r[k][2] = p.x * q.y - p.y * q.x;
                }
                tcache.col[0] = s_col[w][i0]; tcache.col[1] = s_col[w][i1]; tcache.col[2] = s_col[w][i2];
                a.tri_cache[(size_t)slot * 128u + t] = tcache;
            }
            if (cull_faces) {
                const float det = det3_xyw(c0, c1, c2);
                const bool culled = tdet < 0.0f ? det < 0.0f : det > 0.0f;
                if (culled) continue;
            }
            const uint32_t id = (d << VKGV_TRIANGLE_BITS) | t;
            const SVtx v0 = s_sv[w][i0], v1 = s_sv[w][i1], v2 = s_sv[w][i2];
            if (!(v0.ok & v1.ok & v2.ok)) {
                const uint32_t slot = atomicAdd(&a.ctr->clip_count, 1u);
                if (slot < a.clip_cap) {
                    ClipTri ct;
                    ct.c[0] = c0; ct.c[1] = c1; ct.c[2] = c2;
                    ct.id = id;
                    ct.pad[0] = ct.pad[1] = ct.pad[2] = 0u;
                    a.clipq[slot] = ct;
                } else {
                    atomicOr(&a.ctr->overflow, 2u);
                }
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
                push_big(ts, id, a.big, a.big_chunk, a.big_cap, a.ctr);
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

        uint32_t nxt = 0;
        if (lane == 0) nxt = atomicAdd(a.cursor, 1u);
        idx = nwaves + __builtin_amdgcn_readfirstlane(__shfl(nxt, 0));
    }
}

// ---------------------------------------------------------------------------
struct BigArgs {
    BigTri* big;
    uint32_t* big_chunk;
    uint32_t big_cap;
    const ClipTri* clipq;
    uint32_t clip_cap;
    Counters* ctr;
    unsigned long long* keys;
    uint32_t W, H;
    uint32_t tiles_x, tiles_y;
    uint32_t* bin_count;
    uint32_t* bin_start;
    uint32_t* bins;
    uint32_t bin_cap;
    BigSetup* setups;
    BinPair* pairs;
};

// Clipping (rare): Sutherland-Hodgman against pa*x + pb*y + pc*z + pd*w >= 0.
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

// Near plane (z <= w) then the four guard planes; the fan (v0, vi, vi+1) of
// the clipped polygon (at most 3 + 5 vertices) joins the big triangles.
__global__ __launch_bounds__(256) void clip_kernel(BigArgs a)
{
    const uint32_t n = min(a.ctr->clip_count, a.clip_cap);
    const float gx = kClipGuardPx / (float)a.W, gy = kClipGuardPx / (float)a.H;
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const ClipTri ct = a.clipq[i];
        float4 A[10], B[10];
        A[0] = ct.c[0]; A[1] = ct.c[1]; A[2] = ct.c[2];
        int m = 3;
        m = clip_plane(A, m, B, 0.0f, 0.0f, -1.0f, 1.0f);
        m = clip_plane(B, m, A, -1.0f, 0.0f, 0.0f, gx);
        m = clip_plane(A, m, B, 1.0f, 0.0f, 0.0f, gx);
        m = clip_plane(B, m, A, 0.0f, -1.0f, 0.0f, gy);
        m = clip_plane(A, m, B, 0.0f, 1.0f, 0.0f, gy);
        if (m < 3) continue;
        atomicAdd(&a.ctr->clipped, 1u);
        SVtx sv[10];
        bool ok = true;
        for (int k = 0; k < m; ++k) ok = ok && project_vtx(B[k], hw, hh, sv[k]);
        if (!ok) continue;
        for (int k = 1; k + 1 < m; ++k) {
            TriSetup t;
            if (setup_tri(sv[0], sv[k], sv[k + 1], (int)a.W, (int)a.H, t))
                push_big(t, ct.id, a.big, a.big_chunk, a.big_cap, a.ctr);
        }
    }
}

// Edge k of an oriented big triangle: from vertex (k+1)%3 to (k+2)%3.
struct Edge {
    long long ex, ey, ax, ay;
    bool tl;
};

__device__ __forceinline__ void tri_edges(const BigTri& t, Edge* e)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ka = (k + 1) % 3, kb = (k + 2) % 3;
        e[k].ex = (long long)t.x[kb] - t.x[ka];
        e[k].ey = (long long)t.y[kb] - t.y[ka];
        e[k].ax = t.x[ka];
        e[k].ay = t.y[ka];
        e[k].tl = e[k].ey < 0 || (e[k].ey == 0 && e[k].ex > 0);
    }
}

__device__ __forceinline__ long long edge_at(const Edge& e, long long sx, long long sy)
{
    return e.ex * (sy - e.ay) - e.ey * (sx - e.ax);
}

__device__ __forceinline__ uint32_t chunks_of(const BigTri& t)
{
    const int px0 = (int)(t.bb0 & 0xFFFFu), py0 = (int)(t.bb0 >> 16);
    const int px1 = (int)(t.bb1 & 0xFFFFu), py1 = (int)(t.bb1 >> 16);
    const uint32_t ntx = (uint32_t)(px1 / (int)kBinTile - px0 / (int)kBinTile + 1);
    const uint32_t nty = (uint32_t)(py1 / (int)kBinTile - py0 / (int)kBinTile + 1);
    return (ntx * nty + 63u) >> 6;
}

// Chunks of the pass and the triangles they belong to (slots past the queue
// capacity were dropped together with their chunks).
__device__ __forceinline__ void pass_extent(const BigArgs& a, uint32_t& nbig, uint32_t& nchunks)
{
    const unsigned long long packed = a.ctr->big_packed;
    const uint32_t raw = (uint32_t)(packed >> 32);
    nbig = min(raw, a.big_cap);
    nchunks = (uint32_t)packed;
    if (raw > a.big_cap && nbig) nchunks = a.big_chunk[nbig - 1] + chunks_of(a.big[nbig - 1]);
}

// One wave per 64-tile chunk of a big triangle's bounding-box tiles.  The
// wave of chunk 0 stores the triangle's rasterization setup.  Lane l tests
// tile 64*chunk + l: a tile is skipped when one edge is negative at all four
// corner pixel centres of the tile's part of the bounding box (no centre there
// can be covered); kept tiles write pair 64*chunk + l and count themselves.
// The last block to finish scans the per-tile counts.
__global__ __launch_bounds__(256) void bin_kernel(BigArgs a)
{
    __shared__ uint32_t s_scan[4];
    __shared__ bool s_last;
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * 4u + w, nw = gridDim.x * 4u;
    uint32_t nbig, nchunks;
    pass_extent(a, nbig, nchunks);
    if (nchunks == 0) return;   // uniform over the grid: no big triangle this pass
    const bool pair_overflow = (unsigned long long)nchunks * 64ull > a.bin_cap;
    for (uint32_t c = gw; c < nchunks && !pair_overflow; c += nw) {
        uint32_t lo = 0, hi = nbig - 1u;   // last triangle whose first chunk is <= c
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1u) >> 1;
            if (a.big_chunk[mid] <= c) lo = mid; else hi = mid - 1u;
        }
        const uint32_t ti = lo;
        const BigTri t = a.big[ti];
        const uint32_t j = c - a.big_chunk[ti];
        Edge e[3];
        tri_edges(t, e);
        const int px0 = (int)(t.bb0 & 0xFFFFu), py0 = (int)(t.bb0 >> 16);
        const int px1 = (int)(t.bb1 & 0xFFFFu), py1 = (int)(t.bb1 >> 16);
        if (j == 0 && lane == 0) {
            BigSetup s;
            const long long sx0 = (long long)px0 * 256 + 128, sy0 = (long long)py0 * 256 + 128;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                s.e0[k] = edge_at(e[k], sx0, sy0) - (e[k].tl ? 0 : 1);
                s.ex[k] = (int)e[k].ex;
                s.ey[k] = (int)e[k].ey;
            }
            const long long area = ((long long)t.x[1] - t.x[0]) * ((long long)t.y[2] - t.y[0])
                                   - ((long long)t.y[1] - t.y[0]) * ((long long)t.x[2] - t.x[0]);
            depth_plane(t.x, t.y, t.z, area, s.za, s.zb);
            s.z0 = t.z[0];
            s.x0 = t.x[0];
            s.y0 = t.y[0];
            s.id = t.id;
            s.bb0 = t.bb0;
            s.bb1 = t.bb1;
            a.setups[ti] = s;
        }
        const int tx0 = px0 / (int)kBinTile, tx1 = px1 / (int)kBinTile;
        const int ty0 = py0 / (int)kBinTile, ty1 = py1 / (int)kBinTile;
        const int ntx = tx1 - tx0 + 1;
        const uint32_t nt = (uint32_t)(ntx * (ty1 - ty0 + 1));
        const uint32_t k = 64u * j + lane;
        BinPair bp;
        bp.tile = 0xFFFFFFFFu;
        bp.tri = ti;
        if (k < nt) {
            const int ty = ty0 + (int)k / ntx, tx = tx0 + (int)k % ntx;
            const long long cx0 = (long long)max(tx * (int)kBinTile, px0) * 256 + 128;
            const long long cx1 = (long long)min(tx * (int)kBinTile + (int)kBinTile - 1, px1) * 256 + 128;
            const long long cy0 = (long long)max(ty * (int)kBinTile, py0) * 256 + 128;
            const long long cy1 = (long long)min(ty * (int)kBinTile + (int)kBinTile - 1, py1) * 256 + 128;
            bool out = false;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const long long m0 = max(edge_at(e[q], cx0, cy0), edge_at(e[q], cx1, cy0));
                const long long m1 = max(edge_at(e[q], cx0, cy1), edge_at(e[q], cx1, cy1));
                out = out || max(m0, m1) < 0;
            }
            if (!out) {
                bp.tile = (uint32_t)ty * a.tiles_x + (uint32_t)tx;
                atomicAdd(&a.bin_count[bp.tile], 1u);
            }
        }
        a.pairs[64u * c + lane] = bp;
    }

    // last block: exclusive scan of the tile counts
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&a.ctr->bin_done, 1u) == gridDim.x - 1u;
    }
    __syncthreads();
    if (!s_last) return;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);   // the other blocks' counts are visible
    const uint32_t ntiles = a.tiles_x * a.tiles_y;
    const uint32_t per = (ntiles + 255u) / 256u;
    const uint32_t b = threadIdx.x * per, e = min(b + per, ntiles);
    uint32_t sum = 0, most = 0;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t cnt = a.bin_count[i];
        sum += cnt;
        most = max(most, cnt);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) most = max(most, __shfl_xor(most, off));
    if (lane == 0) atomicMax(&a.ctr->max_tile_tris, most);
    uint32_t x = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) s_scan[w] = x;
    __syncthreads();
    uint32_t wbase = 0;
    for (uint32_t q = 0; q < w; ++q) wbase += s_scan[q];
    const uint32_t total = s_scan[0] + s_scan[1] + s_scan[2] + s_scan[3];
    uint32_t run = wbase + x - sum;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t cnt = a.bin_count[i];
        a.bin_start[i] = run;
        a.bin_count[i] = 0u;   // the fill pass counts again (and leaves zero on overflow)
        run += cnt;
    }
    if (threadIdx.x == 0) {
        a.ctr->bin_total = total;
        a.ctr->big_total_tris += nbig;
        a.ctr->big_total_pairs += total;
        a.ctr->bin_done = 0u;
        if (pair_overflow) atomicOr(&a.ctr->overflow, 4u);
    }
}

// Scatter of the pairs into per-tile triangle lists (counts rebuilt).
__global__ __launch_bounds__(256) void bin_fill_kernel(BigArgs a)
{
    uint32_t nbig, nchunks;
    pass_extent(a, nbig, nchunks);
    const unsigned long long np = (unsigned long long)nchunks * 64ull;
    if (np > a.bin_cap) return;   // overflow: this pass's big triangles are dropped
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < (uint32_t)np; i += gridDim.x * 256u) {
        const BinPair p = a.pairs[i];
        if (p.tile == 0xFFFFFFFFu) continue;
        const uint32_t slot = a.bin_start[p.tile] + atomicAdd(&a.bin_count[p.tile], 1u);
        a.bins[slot] = p.tri;
    }
}

// Up to kTileSplit blocks per 32x32 screen tile, one per kPerBlock triangles
// of its list: block s of S rasterizes triangles s, s + S, ... into LDS keys,
// then merges its tile into the visibility buffer.  Setups are staged through
// LDS kChunk at a time.  Thread t owns pixels (t & 31, (t >> 5) + 8q), q = 0..3.
__global__ __launch_bounds__(256) void tile_kernel(BigArgs a)
{
    constexpr uint32_t kChunk = 32;
    constexpr uint32_t kPerBlock = 32;
    constexpr uint32_t kWords = sizeof(BigSetup) / 4;
    __shared__ unsigned long long s_key[kBinTile * kBinTile];
    __shared__ uint32_t s_idx[kChunk];
    __shared__ BigSetup s_set[kChunk];
    const uint32_t tile = blockIdx.x / kTileSplit, part = blockIdx.x % kTileSplit;
    const uint32_t n_all = a.bin_count[tile];
    const uint32_t split = min(kTileSplit, (n_all + kPerBlock - 1u) / kPerBlock);
    if (part >= split) return;
    const uint32_t n = (n_all - part + split - 1u) / split;   // this block's share
    const uint32_t start = a.bin_start[tile];
    const int ox = (int)((tile % a.tiles_x) * kBinTile), oy = (int)((tile / a.tiles_x) * kBinTile);
    const int lx = (int)(threadIdx.x & 31u), ly = (int)(threadIdx.x >> 5);
    const int px = ox + lx, py_first = oy + ly;
#pragma unroll
    for (int q = 0; q < 4; ++q) s_key[threadIdx.x + 256 * q] = VKGV_KEY_CLEAR;
    const int sx = px * 256 + 128;
    for (uint32_t c0 = 0; c0 < n; c0 += kChunk) {
        const uint32_t m = min(kChunk, n - c0);
        if (threadIdx.x < m) s_idx[threadIdx.x] = a.bins[start + part + (c0 + threadIdx.x) * split];
        __syncthreads();
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_set);
        for (uint32_t wd = threadIdx.x; wd < m * kWords; wd += 256u)
            dst[wd] = reinterpret_cast<const uint32_t*>(a.setups + s_idx[wd / kWords])[wd % kWords];
        __syncthreads();
        for (uint32_t k = 0; k < m; ++k) {
            const BigSetup& s = s_set[k];
            const int bx0 = (int)(s.bb0 & 0xFFFFu), by0 = (int)(s.bb0 >> 16);
            const int bx1 = (int)(s.bb1 & 0xFFFFu), by1 = (int)(s.bb1 >> 16);
            if (px < bx0 || px > bx1) continue;
            long long ev[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                ev[r] = s.e0[r] - (long long)(px - bx0) * (long long)s.ey[r] * 256
                        + (long long)(py_first - by0) * (long long)s.ex[r] * 256;
            const float dxf = (float)(sx - s.x0) * 0.00390625f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int py = py_first + 8 * q;
                if (q) {
#pragma unroll
                    for (int r = 0; r < 3; ++r) ev[r] += (long long)s.ex[r] * 2048;   // +8 rows
                }
                if (py < by0 || py > by1) continue;
                if ((ev[0] | ev[1] | ev[2]) < 0) continue;
                const float dyf = (float)(py * 256 + 128 - s.y0) * 0.00390625f;
                const float z = s.z0 + (s.za * dxf + s.zb * dyf);
                if (!(z >= 0.0f && z <= 1.0f)) continue;
                (void)__hip_atomic_fetch_min(&s_key[threadIdx.x + 256 * q], make_key(z, s.id), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        __syncthreads();
    }
    if (px >= (int)a.W) return;
    // A tile owned by one block has no concurrent writer (the meshlet raster
    // of this pass has finished; tiles are disjoint): plain read-min-store.
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int py = py_first + 8 * q;
        const unsigned long long key = s_key[threadIdx.x + 256 * q];
        if (key == VKGV_KEY_CLEAR || py >= (int)a.H) continue;
        unsigned long long* p = a.keys + (size_t)py * a.W + px;
        if (key < *p) {
            if (split == 1) *p = key;
            else key_min(p, key);
        }
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
    a.big_chunk = v.big_chunk;
    a.big_cap = v.big_capacity;
    a.clipq = v.clipq;
    a.clip_cap = v.clip_capacity;
    a.ctr = v.counters;
    a.slot_base = list ? &v.counters->vis_count[0] : nullptr;
    a.tri_cache = v.tri_cache;
    a.slot_albedo = v.slot_albedo;
    a.draw_slot = v.draw_slot;
    a.cache_slots = v.cache_slots;
    raster_kernel<<<grid_blocks, kBlock, 0, st>>>(a);
}

void launch_raster_big(const ViewDev& v, hipStream_t st)
{
    BigArgs a;
    a.big = v.big;
    a.big_chunk = v.big_chunk;
    a.big_cap = v.big_capacity;
    a.clipq = v.clipq;
    a.clip_cap = v.clip_capacity;
    a.ctr = v.counters;
    a.keys = v.keys;
    a.W = v.width;
    a.H = v.height;
    a.tiles_x = v.tiles_x;
    a.tiles_y = v.tiles_y;
    a.bin_count = v.bin_count;
    a.bin_start = v.bin_start;
    a.bins = v.bins;
    a.bin_cap = v.bin_capacity;
    a.setups = v.setups;
    a.pairs = v.pairs;
    clip_kernel<<<16, 256, 0, st>>>(a);
    bin_kernel<<<1024, 256, 0, st>>>(a);
    bin_fill_kernel<<<256, 256, 0, st>>>(a);
    tile_kernel<<<v.tiles_x * v.tiles_y * kTileSplit, 256, 0, st>>>(a);
}

} // namespace vkgv
