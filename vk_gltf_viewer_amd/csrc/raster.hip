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
//  raster_kernel  persistent grid of 4-wave blocks; wave g takes visible
//                 meshlets g, g + waves, ..., transforms its <= 64 vertices into
//                 LDS, back-face culls its <= 124 triangles (2 per lane) and
//                 rasterizes the small ones cooperatively: a wave prefix sum
//                 over their bounding-box pixel counts gives each lane an equal
//                 run of pixel slots, walked with integer edge steps (a heavy
//                 meshlet hands most of its slots to the bin pass as deferred
//                 runs).  Large triangles go to the big queue (a slot, and the
//                 64-tile chunks of the triangle's 32x32-pixel bounding-box
//                 tiles in the chunk list), triangles needing clipping to the
//                 clip queue.  A
//                 meshlet with a triangle that found either full is listed for
//                 the direct path.
//  clip_kernel    clips the queued triangles, one per wave; their fans join the
//                 big queue (or, without room, the direct path).  Then redoes the listed
//                 meshlets: each of their triangles that is not small takes the
//                 direct path, where the wave sweeps one triangle's bounding
//                 box at a time, so that no triangle is lost when a queue is
//                 full (slower, same result).
//  bin_kernel     one wave per chunk: stores the triangle's setup (chunk 0),
//                 tests 64 tiles and appends the triangle to the fixed-size
//                 list of every tile that can hold a covered pixel centre (a
//                 tile whose list is full gets the triangle swept directly).
//                 Then the deferred fragment runs of heavy meshlets.
//  tile_kernel    static work items (tile, 32x16 band, part of the tile's
//                 list), each rasterizing its share of the tile's triangles
//                 into keys held in registers, then merging them into the
//                 visibility buffer.
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

// A small triangle staged in LDS (52 bytes).  The edge steps A = -ey * 256
// and B = ex * 256 are kept as the 16-bit deltas (the small path's extents
// stay below 2^15 sub-pixel units), and the depth-plane origin offsets as the
// 16-bit sub-pixel distances they are; both are widened exactly on load.
struct SmallTri {
    int C[3];
    float za, zb, z0;
    uint32_t org;   // px0 | py0 << 16
    uint32_t id;
    int16_t ex[3], ey[3];
    int16_t ox, oy; // sub-pixel offsets of the bbox origin's centre from vertex 0
    uint16_t dim;   // bw | bh << 8
    uint16_t pad;
};
static_assert(sizeof(SmallTri) == 52, "SmallTri layout");

// The register form of a staged small triangle.
struct SmallTriR {
    int C[3], A[3], B[3];
    float za, zb, z0, oxf, oyf;
    uint32_t org, id, bw;
};

__device__ __forceinline__ SmallTriR load_small(const SmallTri& s)
{
    SmallTriR r;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r.C[k] = s.C[k];
        r.A[k] = -(int)s.ey[k] * 256;
        r.B[k] = (int)s.ex[k] * 256;
    }
    r.za = s.za; r.zb = s.zb; r.z0 = s.z0;
    r.oxf = (float)s.ox * 0.00390625f;
    r.oyf = (float)s.oy * 0.00390625f;
    r.org = s.org;
    r.id = s.id;
    r.bw = s.dim & 0xFFu;
    return r;
}

struct RasterArgs {
    SceneDev s;
    const RasterRec* list;
    unsigned long long* keys;
    uint32_t W, H;
    uint32_t flags;
    int small_lim;                  // small path iff bbox extent < small_lim (1/256 pixel units)
    BigTri* big;
    uint32_t* big_chunk;
    uint32_t big_cap;
    uint32_t* chunk_tri;
    uint32_t chunk_cap;
    ClipTri* clipq;
    uint32_t clip_cap;
    Counters* ctr;
    uint32_t* redo;
    const float* mvp;               // the view's VP M table (kMvpStride floats per transform)
    uint32_t count;                 // meshlets in the list
    uint32_t first;                 // the wave's first meshlet of the list (then every nwaves-th)
    uint32_t slot0;                 // first resolve-cache slot of this list
    TriCache* tri_cache;
    float4* slot_albedo;
    uint32_t* slot_mat;
    uint32_t* draw_slot;
    uint32_t cache_slots;
    FragStage* stage;               // deferred fragment runs of heavy meshlets
    FragJob* jobs;
    uint32_t stage_cap, job_cap;
};

// A wave-uniform value read by a vector load, moved to a scalar register.
__device__ __forceinline__ uint32_t uni(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)x); }
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T>
__device__ __forceinline__ T* uni(T* p)
{
    const unsigned long long u = (unsigned long long)p;
    const uint32_t lo = uni((uint32_t)u), hi = uni((uint32_t)(u >> 32));
    return (T*)(((unsigned long long)hi << 32) | lo);
}

// Loads and stores through global (address space 1) pointers.  A pointer
// read from memory is generic to the compiler, and a generic (flat) access
// also counts in the LDS counter, so every later LDS wait of the wave would
// wait for it; a global one is counted with vector memory only.
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, const T& v)
{
    *(__attribute__((address_space(1))) T*)p = v;
}
// 16- and 8-byte records through the compiler's own vector types (HIP's
// vector classes have no assignment to an address-space-qualified object)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void gstore(uint4* p, const uint4& v)
{
    const u32x4 x = { v.x, v.y, v.z, v.w };
    *(__attribute__((address_space(1))) u32x4*)p = x;
}
__device__ __forceinline__ void gstore(float4* p, const float4& v)
{
    gstore(reinterpret_cast<uint4*>(p),
           make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w)));
}
__device__ __forceinline__ void gstore(FragJob* p, const FragJob& v)
{
    const u32x2 x = { v.stage, v.first };
    *(__attribute__((address_space(1))) u32x2*)p = x;
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
// Issued through a global (address space 1) pointer: a flat atomic would also
// count in the LDS counter, so the raster's next LDS read would wait for
// every atomic in flight; a global one waits for none of them.
typedef __attribute__((address_space(1))) unsigned long long global_u64;
__device__ __forceinline__ void key_min(unsigned long long* p, unsigned long long key)
{
    (void)__hip_atomic_fetch_min((global_u64*)p, key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// No early depth read: the wave does not wait on memory inside the pixel loop.
__device__ __forceinline__ void write_fragment(unsigned long long* keys, uint32_t W, int px, int py, float z, uint32_t id)
{
    key_min(keys + (size_t)py * W + px, make_key(z, id));
}

// Per-lane raster statistics of frames with VKGV_STATS: triangles set up,
// fragments (covered centres with 0 <= z <= 1), and how their visibility
// updates reached memory (64-bit atomicMin, or a tile's plain
// read-min-store).  The STATS = false instantiations count nothing.
template <bool STATS>
struct RasterCount {
    uint32_t tris = 0, frags = 0, atomics = 0, stores = 0;
    __device__ __forceinline__ void tri() { if constexpr (STATS) ++tris; }
    __device__ __forceinline__ void frag_atomic() { if constexpr (STATS) { ++frags; ++atomics; } }
    __device__ __forceinline__ void frag() { if constexpr (STATS) ++frags; }
    __device__ __forceinline__ void atomic() { if constexpr (STATS) ++atomics; }
    __device__ __forceinline__ void store() { if constexpr (STATS) ++stores; }
    // Wave sums into the view's counters (all 64 lanes call it), then zero.
    __device__ __forceinline__ void flush(Counters* c)
    {
        if constexpr (STATS) {
            unsigned long long v[4] = { tris, frags, atomics, stores };
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
            if ((threadIdx.x & 63u) == 0) {
                if (v[0]) atomicAdd(&c->raster_tris, v[0]);
                if (v[1]) atomicAdd(&c->fragments, v[1]);
                if (v[2]) atomicAdd(&c->frag_atomics, v[2]);
                if (v[3]) atomicAdd(&c->frag_stores, v[3]);
            }
            tris = frags = atomics = stores = 0;
        }
    }
};

// Edge k of an oriented big triangle: from vertex (k+1)%3 to (k+2)%3.
struct Edge {
    long long ex, ey, ax, ay;
    bool tl;
};

__device__ __forceinline__ void edges_of(const int* x, const int* y, Edge* e)
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int ka = (k + 1) % 3, kb = (k + 2) % 3;
        e[k].ex = (long long)x[kb] - x[ka];
        e[k].ey = (long long)y[kb] - y[ka];
        e[k].ax = x[ka];
        e[k].ay = y[ka];
        e[k].tl = e[k].ey < 0 || (e[k].ey == 0 && e[k].ex > 0);
    }
}

__device__ __forceinline__ long long edge_at(const Edge& e, long long sx, long long sy)
{
    return e.ex * (sy - e.ay) - e.ey * (sx - e.ax);
}

// A whole wave sweeps the pixel rectangle [rx0, rx1] x [ry0, ry1] of one
// triangle (wave-uniform arguments: edges, snapped vertices, depth plane) in
// 8x8-pixel blocks with the tile pass's coverage rule and depth expression,
// writing its fragments with atomicMin, so it writes exactly the fragments
// the tile pass would.
template <bool STATS>
__device__ void sweep_rect(const Edge* e, const int* x, const int* y, const float* z, float za, float zb, uint32_t id,
                           int rx0, int ry0, int rx1, int ry1, unsigned long long* keys, uint32_t W,
                           RasterCount<STATS>& rc)
{
    const uint32_t lane = threadIdx.x & 63u;
    const int lx = (int)(lane & 7u), ly = (int)(lane >> 3);
    for (int by = ry0; by <= ry1; by += 8) {
        const int py = by + ly;
        const long long sy = (long long)py * 256 + 128;
        const float dyf = (float)(py * 256 + 128 - y[0]) * 0.00390625f;
        for (int bx = rx0; bx <= rx1; bx += 8) {
            const int px = bx + lx;
            if (px > rx1 || py > ry1) continue;
            const long long sx = (long long)px * 256 + 128;
            bool in = true;
#pragma unroll
            for (int q = 0; q < 3; ++q) in = in && edge_at(e[q], sx, sy) - (e[q].tl ? 0 : 1) >= 0;
            if (!in) continue;
            const float dxf = (float)(px * 256 + 128 - x[0]) * 0.00390625f;
            const float zz = z[0] + (za * dxf + zb * dyf);
            if (!(zz >= 0.0f && zz <= 1.0f)) continue;
            write_fragment(keys, W, px, py, zz, id);
            rc.frag_atomic();
        }
    }
}

// Triangles that found a queue full: every lane with `want` set hands its
// triangle to the whole wave, one triangle at a time, and the wave sweeps its
// bounding box (sweep_rect).  Called by all 64 lanes of a wave.  Slower than
// the binned path, but no triangle is lost.
template <bool STATS>
__device__ void wave_sweep(bool want, const TriSetup& t, uint32_t id, unsigned long long* keys, uint32_t W,
                           RasterCount<STATS>& rc)
{
    unsigned long long pending = __ballot(want);
    while (pending != 0ull) {
        const int j = __ffsll((long long)pending) - 1;
        pending &= pending - 1ull;
        int x[3], y[3];
        float z[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            x[k] = __shfl(t.x[k], j);
            y[k] = __shfl(t.y[k], j);
            z[k] = __shfl(t.z[k], j);
        }
        const int px0 = __shfl(t.px0, j), px1 = __shfl(t.px1, j);
        const int py0 = __shfl(t.py0, j), py1 = __shfl(t.py1, j);
        const long long area = __shfl(t.area, j);
        const uint32_t tid = __shfl(id, j);
        Edge e[3];
        edges_of(x, y, e);
        float za, zb;
        depth_plane(x, y, z, area, za, zb);
        sweep_rect(e, x, y, z, za, zb, tid, px0, py0, px1, py1, keys, W, rc);
    }
}

// 64-tile chunks a big triangle reserves in the chunk list: none when it
// needs more than the whole list (it is rasterized directly).
__device__ __forceinline__ uint32_t reserve_chunks(const TriSetup& t, uint32_t chunk_cap)
{
    const uint32_t ntx = (uint32_t)(t.px1 / (int)kBinTile - t.px0 / (int)kBinTile + 1);
    const uint32_t nty = (uint32_t)(t.py1 / (int)kBinTile - t.py0 / (int)kBinTile + 1);
    const uint32_t c = (ntx * nty + 63u) >> 6;
    return c <= chunk_cap ? c : 0u;
}

// Where a big triangle goes, and what it needs to get there.
struct BigQueue {
    BigTri* big;
    uint32_t* big_chunk;
    uint32_t cap;
    uint32_t* chunk_tri;
    uint32_t chunk_cap;
    Counters* ctr;
    unsigned long long* keys;
    uint32_t W;
};

// Stores a big triangle in `slot` with its chunks from c0 (chunk_tri marks
// them as the triangle's) and returns true.  Without room in the queue or in
// the chunk list it returns false, marks its slot empty (big_chunk ~0u; its
// reserved chunks keep stale chunk_tri entries, which the bin pass rejects)
// and sets overflow bit 0 (queue) or 2 (chunk list): the frame takes the
// slower direct path for the triangle.
__device__ __forceinline__ bool store_big(const TriSetup& t, uint32_t id, uint32_t slot, unsigned long long c0,
                                          uint32_t chunks, const BigQueue& q)
{
    if (chunks == 0u || slot >= q.cap || c0 + chunks > q.chunk_cap) {
        if (slot < q.cap) q.big_chunk[slot] = 0xFFFFFFFFu;
        atomicOr(&q.ctr->overflow, slot >= q.cap ? 1u : 4u);
        return false;
    }
    q.big_chunk[slot] = (uint32_t)c0;
    for (uint32_t k = 0; k < chunks; ++k) q.chunk_tri[c0 + k] = slot;
    BigTri b;
#pragma unroll
    for (int k = 0; k < 3; ++k) { b.x[k] = t.x[k]; b.y[k] = t.y[k]; b.z[k] = t.z[k]; }
    b.id = id;
    b.bb0 = (uint32_t)t.px0 | ((uint32_t)t.py0 << 16);
    b.bb1 = (uint32_t)t.px1 | ((uint32_t)t.py1 << 16);
    q.big[slot] = b;
    return true;
}

// Appends one big triangle (one lane): a slot and its chunks.  Returns false
// when there was no room.
__device__ __forceinline__ bool push_big(const TriSetup& t, uint32_t id, const BigQueue& q)
{
    const uint32_t chunks = reserve_chunks(t, q.chunk_cap);
    const uint32_t slot = atomicAdd(&q.ctr->big_count, 1u);
    const unsigned long long c0 = chunks ? atomicAdd(&q.ctr->big_chunks, (unsigned long long)chunks) : 0ull;
    return store_big(t, id, slot, c0, chunks, q);
}

// Wave-wide append (called by all 64 lanes): a prefix sum of the lanes'
// (1, chunks) pairs and one atomic per counter and wave reserve consecutive
// slots and chunks in lane order.  A lane's chunk count is at most the chunk
// list's chunk capacity (below 2^19), so the 64-lane sums fit their 32-bit
// halves.  Returns true on a lane whose triangle found no room.
__device__ __forceinline__ bool push_big_wave(bool want, const TriSetup& t, uint32_t id, uint32_t lane,
                                              const BigQueue& q)
{
    const uint32_t chunks = want ? reserve_chunks(t, q.chunk_cap) : 0u;
    const unsigned long long mine = want ? ((1ull << 32) | (unsigned long long)chunks) : 0ull;
    unsigned long long x = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    const unsigned long long total = __shfl(x, 63);
    if (total == 0ull) return false;   // wave-uniform
    // slot base << 32 | chunk base (saturated: a base past the list fails either way)
    unsigned long long base = 0ull;
    if (lane == 0) {
        const uint32_t sb = atomicAdd(&q.ctr->big_count, (uint32_t)(total >> 32));
        const unsigned long long cb =
            (uint32_t)total ? atomicAdd(&q.ctr->big_chunks, (unsigned long long)(uint32_t)total) : 0ull;
        base = ((unsigned long long)sb << 32) | (cb < 0xFFFFFFFFull ? cb : 0xFFFFFFFFull);
    }
    base = __shfl(base, 0);
    const unsigned long long before = x - mine;
    return want && !store_big(t, id, (uint32_t)(base >> 32) + (uint32_t)(before >> 32),
                              (base & 0xFFFFFFFFull) + (uint32_t)before, chunks, q);
}

// Clipping (rare): Sutherland-Hodgman against pa*x + pb*y + pc*z + pd*w >= 0.
// A convex polygon gains at most one vertex per plane (3 + 5 <= kClipMax);
// non-finite or degenerate input can produce more crossings: such a polygon
// is dropped (returns 0) and nothing is written past kClipMax.
constexpr int kClipMax = 16;
constexpr uint32_t kClipThreads = 64;   // one wave per clip block

// A polygon of one lane: vertex k at base[k] (private memory or LDS).
struct LocalPoly {
    float4* base;
    __device__ __forceinline__ float4& operator[](int k) const { return base[k]; }
};

template <typename Poly>
__device__ __forceinline__ int clip_plane(const Poly in, int n, const Poly out, float pa, float pb, float pc, float pd)
{
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const float4 p = in[i];
        const float4 q = in[i + 1 < n ? i + 1 : 0];
        const float dp = ((pa * p.x + pb * p.y) + pc * p.z) + pd * p.w;
        const float dq = ((pa * q.x + pb * q.y) + pc * q.z) + pd * q.w;
        if (dp >= 0.0f) {
            if (m < kClipMax) out[m] = p;
            ++m;
        }
        if ((dp >= 0.0f) != (dq >= 0.0f)) {
            const float t = dp / (dp - dq);
            float4 r;
            r.x = p.x + t * (q.x - p.x);
            r.y = p.y + t * (q.y - p.y);
            r.z = p.z + t * (q.z - p.z);
            r.w = p.w + t * (q.w - p.w);
            if (m < kClipMax) out[m] = r;
            ++m;
        }
    }
    return m <= kClipMax ? m : 0;
}

// Near plane (z <= w) then the four guard planes.  pa and pb are the lane's
// two polygon buffers; the clipped polygon (at most 3 + 5 vertices) ends in
// pb.  Returns its vertex count, or 0 when fewer than 3 remain.
template <typename Poly>
__device__ __forceinline__ int clip_polygon(const Poly pa, const Poly pb, float4 c0, float4 c1, float4 c2, uint32_t W,
                                            uint32_t H)
{
    const float gx = kClipGuardPx / (float)W, gy = kClipGuardPx / (float)H;
    pa[0] = c0; pa[1] = c1; pa[2] = c2;
    int m = 3;
    m = clip_plane(pa, m, pb, 0.0f, 0.0f, -1.0f, 1.0f);
    m = clip_plane(pb, m, pa, -1.0f, 0.0f, 0.0f, gx);
    m = clip_plane(pa, m, pb, 1.0f, 0.0f, 0.0f, gx);
    m = clip_plane(pb, m, pa, 0.0f, -1.0f, 0.0f, gy);
    m = clip_plane(pa, m, pb, 0.0f, 1.0f, 0.0f, gy);
    return m >= 3 ? m : 0;
}

// Whether every vertex of the clipped polygon projects (w > 0, inside the
// guard band); a polygon that does not is dropped.
template <typename Poly>
__device__ __forceinline__ bool polygon_projects(const Poly p, int m, float hw, float hh)
{
    bool ok = true;
    for (int k = 0; k < m && ok; ++k) {
        SVtx sv;
        ok = project_vtx(p[k], hw, hh, sv);
    }
    return ok;
}

// Setup of fan triangle (v0, vf, vf+1) of a projecting polygon.
template <typename Poly>
__device__ __forceinline__ bool fan_setup(const Poly p, int f, float hw, float hh, int W, int H, TriSetup& t)
{
    SVtx s0, s1, s2;
    project_vtx(p[0], hw, hh, s0);
    project_vtx(p[f], hw, hh, s1);
    project_vtx(p[f + 1], hw, hh, s2);
    return setup_tri(s0, s1, s2, W, H, t);
}

__device__ __forceinline__ uint32_t tri_byte(const uint32_t* words, uint32_t b)
{
    return (words[b >> 2] >> ((b & 3u) * 8u)) & 0xFFu;
}

// Words of a FragStage record: the staged triangles, then the 128-entry slot
// prefix, then the slot total.
constexpr uint32_t kStageTriWords = VKGV_MESHLET_TRIANGLES * (uint32_t)(sizeof(SmallTri) / 4);
constexpr uint32_t kStageTri4 = kStageTriWords / 4;                 // the triangles as 16-byte words
static_assert(kStageTriWords % 4 == 0, "the staged triangles are copied 16 bytes at a time");
static_assert(kStageTriWords + 128u + 4u <= kFragStageWords, "a staged meshlet fits its record");

// Pixel slots [k, kend) of a meshlet's small triangles (tri[] staged setups,
// pre[] the 128-entry exclusive prefix of their bounding-box pixel counts,
// `total` their sum), walked by one lane: one search finds the triangle of
// the first slot, then the lane walks the run with integer edge steps and
// moves to the next triangle when the run crosses a triangle's start (no
// search per pixel).  Covered centres with 0 <= z <= 1 update the
// visibility buffer with a 64-bit atomicMin.
template <bool STATS>
__device__ __forceinline__ void walk_small(const SmallTri* tri, const uint32_t* pre, uint32_t total, uint32_t k,
                                           uint32_t kend, unsigned long long* keys, uint32_t W, RasterCount<STATS>& rc)
{
    if (k >= kend) return;
    uint32_t lo = 0, hi = 127;   // last triangle whose first slot is <= k
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (pre[mid] <= k) lo = mid; else hi = mid - 1u;
    }
    uint32_t next = lo < 127u ? pre[lo + 1u] : total;
    SmallTriR st = load_small(tri[lo]);
    uint32_t bw = st.bw;
    const uint32_t r = k - pre[lo];
    int jj = (int)(r / bw), ii = (int)(r - (uint32_t)jj * bw);
    int e0 = st.C[0] + st.A[0] * ii + st.B[0] * jj;
    int e1 = st.C[1] + st.A[1] * ii + st.B[1] * jj;
    int e2 = st.C[2] + st.A[2] * ii + st.B[2] * jj;
    for (;;) {
        if ((e0 | e1 | e2) >= 0) {
            const float z = st.z0 + (st.za * (st.oxf + (float)ii) + st.zb * (st.oyf + (float)jj));
            if (z >= 0.0f && z <= 1.0f) {
                write_fragment(keys, W, (int)(st.org & 0xFFFFu) + ii, (int)(st.org >> 16) + jj, z, st.id);
                rc.frag_atomic();
            }
        }
        if (++k >= kend) break;
        if (k == next) {   // first slot of a later triangle (empty ones skipped)
            while (lo < 127u && pre[lo + 1u] <= k) ++lo;
            next = lo < 127u ? pre[lo + 1u] : total;
            st = load_small(tri[lo]);
            bw = st.bw;
            ii = 0;
            jj = 0;
            e0 = st.C[0];
            e1 = st.C[1];
            e2 = st.C[2];
        } else if (++ii == (int)bw) {   // next row of the bounding box
            ii = 0;
            ++jj;
            e0 = st.C[0] + st.B[0] * jj;
            e1 = st.C[1] + st.B[1] * jj;
            e2 = st.C[2] + st.B[2] * jj;
        } else {
            e0 += st.A[0];
            e1 += st.A[1];
            e2 += st.A[2];
        }
    }
}

// One view's share of the meshlet raster (wave w of the grid takes meshlets
// w, w + waves, ... of the view's list).
template <bool STATS>
__device__ __forceinline__ void raster_view(const RasterArgs& a)
{
    RasterCount<STATS> rc;
    __shared__ float4 s_clip[kWaves][64];
    __shared__ SVtx s_sv[kWaves][64];
    __shared__ uint32_t s_tri[kWaves][96];
    __shared__ __align__(16) SmallTri s_st[kWaves][VKGV_MESHLET_TRIANGLES];
    __shared__ __align__(16) uint32_t s_pre[kWaves][128];
    __shared__ uint32_t s_list[kWaves][128];   // the meshlet's triangles past the per-slot tests

    const int w = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    __shared__ uint32_t s_col[kWaves][64];
    __shared__ uint32_t s_uv[kWaves][64];
    const uint32_t count = a.count;
    const uint32_t slot0 = a.slot0;
    const uint32_t nwaves = gridDim.x * kWaves;
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    const int W = (int)a.W, H = (int)a.H;
    const BigQueue bq{ a.big, a.big_chunk, a.big_cap, a.chunk_tri, a.chunk_cap, a.ctr, a.keys, a.W };
    // non-zero when a triangle of the wave's meshlet found a queue full
    __shared__ uint32_t s_spill[kWaves];
    if (lane == 0) s_spill[w] = 0u;

    // A software pipeline over the wave's meshlets i = first, first + nwaves,
    // ...: each input of a meshlet is loaded while earlier meshlets are
    // rasterized, so that a meshlet waits on memory once (the drain at its
    // start, which also covers the previous meshlet's fragment atomics) instead
    // of once per dependent load (record -> vertex indices -> vertices):
    //   during meshlet i:  record of i + 3       (lane k < 8: word k)
    //                      vertex indices of i + 2 (lane l < vertex count)
    //                      vertices, VP M row (lane k < 17: float k) and
    //                      triangle words (lanes l and l + 64) of i + 1
    // Records and rows are held one word per lane and read with readlane.
    const uint32_t* listw = reinterpret_cast<const uint32_t*>(a.list);
    const uint32_t* vidx_base = a.s.vidx_base;
    const uint8_t* vtx_bytes = reinterpret_cast<const uint8_t*>(a.s.vtx_base);
    const bool tex = (a.flags & VKGV_RESOLVE_TEXTURED) != 0;   // the textured resolve reads the uvs too
    uint32_t idx = __builtin_amdgcn_readfirstlane(a.first);
    // record word `lane` of list entry j (0 past the list)
    auto rec_word = [&](uint32_t j) -> uint32_t {
        return (j < count && lane < 8u) ? gload(listw + (size_t)j * 8u + lane) : 0u;
    };
    auto field = [](uint32_t r, int k) -> uint32_t { return (uint32_t)__builtin_amdgcn_readlane((int)r, k); };
    // vertex index of lane `lane` of the meshlet with record r
    auto vidx_of = [&](uint32_t r, bool live) -> uint32_t {
        const uint32_t vc = field(r, 6) & 0xFFu;
        return (live && lane < vc) ? gload(vidx_base + field(r, 2) + lane) : 0u;
    };
    struct Fetch {
        float x, y, z;      // vertex position
        uint32_t col, uv;   // vertex colour and uv words
        uint32_t row;       // VP M row word (lanes 0..16)
        uint32_t t0, t1;    // triangle words lane, lane + 64
        uint32_t mat;       // material albedo factor word (lanes 0..3)
    };
    // the inputs of the meshlet with record r and vertex indices vi
    auto fetch = [&](uint32_t r, uint32_t vi, bool live, Fetch& f) {
        f.x = f.y = f.z = 0.0f;
        f.col = f.uv = f.row = f.t0 = f.t1 = f.mat = 0u;
        if (!live) return;
        const uint32_t vc = field(r, 6) & 0xFFu, tc = (field(r, 6) >> 8) & 0xFFu;
        if (lane < vc) {
            const float* pp = reinterpret_cast<const float*>(vtx_bytes + field(r, 4)) + (size_t)vi * 6;
            f.x = gload(pp);
            f.y = gload(pp + 1);
            f.z = gload(pp + 2);
            f.col = gload(reinterpret_cast<const uint32_t*>(pp) + 3);
            if (tex) f.uv = gload(reinterpret_cast<const uint32_t*>(pp) + 5);
        }
        if (lane < 17u) f.row = __float_as_uint(gload(a.mvp + (size_t)field(r, 1) * kMvpStride + lane));
        if (lane < 4u) f.mat = __float_as_uint(gload(&a.s.materials[field(r, 5) & 0x7FFFFFFFu].albedo_factor[lane]));
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.s.tri_base + field(r, 3));
        const uint32_t ndw = (tc * 3u + 3u) >> 2;
        if (lane < ndw) f.t0 = gload(src + lane);
        if (lane + 64u < ndw) f.t1 = gload(src + lane + 64u);
    };
    uint32_t r0 = rec_word(idx), r1 = rec_word(idx + nwaves), r2 = rec_word(idx + 2u * nwaves);
    uint32_t x1 = vidx_of(r1, idx + nwaves < count);
    Fetch f0;
    fetch(r0, vidx_of(r0, idx < count), idx < count, f0);
    while (idx < count) {
        const uint32_t d = field(r0, 0);
        const uint32_t matw = field(r0, 5);
        const uint32_t counts = field(r0, 6);
        const uint32_t nidx = idx + nwaves;
        // VP M and det(M) of the transform, computed once per frame by the first cull
        Mat4 mvp;
#pragma unroll
        for (int k = 0; k < 16; ++k) mvp.m[k] = __uint_as_float(field(f0.row, k));
        const float tdet = __uint_as_float(field(f0.row, 16));
        const bool cull_faces = !(matw >> 31) && !(a.flags & VKGV_NO_BACKFACE);
        // the flat resolve reads only the slot albedo; the interpolating ones the triangles too
        const bool bary = (a.flags & (VKGV_RESOLVE_BARY | VKGV_RESOLVE_TEXTURED | VKGV_MOTION_VECTORS)) != 0;
        const uint32_t vc = counts & 0xFFu, tc = (counts >> 8) & 0xFFu;
        // resolve-cache slot of this meshlet (none when the cache is full)
        const uint32_t slot = slot0 + idx;
        const bool cached = slot < a.cache_slots;
        {
            const float4 albedo = make_float4(__uint_as_float(field(f0.mat, 0)), __uint_as_float(field(f0.mat, 1)),
                                              __uint_as_float(field(f0.mat, 2)), __uint_as_float(field(f0.mat, 3)));
            if (lane == 0) {
                gstore(a.draw_slot + d, cached ? slot : 0xFFFFFFFFu);
                if (cached) {
                    gstore(a.slot_albedo + slot, albedo);
                    gstore(a.slot_mat + slot, matw & 0x7FFFFFFFu);
                }
            }
        }

        // vertices: clip = (VP * M) * (pos, 1) (visbuffer.mesh.glsl:44-65)
        if (lane < vc) {
            const float4 c = mat4_point(mvp, f0.x, f0.y, f0.z);
            s_clip[w][lane] = c;
            s_col[w][lane] = f0.col;
            if (tex) s_uv[w][lane] = f0.uv;
            SVtx sv;
            sv.x = 0; sv.y = 0; sv.z = 0.0f;
            sv.ok = project_vtx(c, hw, hh, sv) ? 1u : 0u;
            s_sv[w][lane] = sv;
        }
        {
            const uint32_t ndw = (tc * 3u + 3u) >> 2;
            if (lane < ndw) s_tri[w][lane] = f0.t0;
            if (lane + 64u < ndw) s_tri[w][lane + 64u] = f0.t1;
        }
        // the pipeline's loads for the next meshlets
        fetch(r1, x1, nidx < count, f0);
        x1 = vidx_of(r2, nidx + nwaves < count);
        r0 = r1;
        r1 = r2;
        r2 = rec_word(nidx + 2u * nwaves);
        wave_sync();

        // Stage 1, every triangle slot: the resolve cache, the back-face test
        // and setup_tri's pixel-range test (a projected triangle whose
        // bounding box holds no pixel centre rasterizes nothing).  The
        // triangles left (and those that need clipping) are packed into
        // s_list (t | i0 << 8 | i1 << 16 | i2 << 24), so that the setup below
        // runs over them in as few 64-lane rounds as there are: one for a
        // typical meshlet, whose back faces and pixel-less triangles are most
        // of its slots.
        uint32_t nlive = 0;
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            const uint32_t t = lane + 64u * (uint32_t)sl;
            bool live = false;
            uint32_t packed = 0u;
            if (t < tc) {
                const uint32_t i0 = tri_byte(s_tri[w], 3u * t), i1 = tri_byte(s_tri[w], 3u * t + 1u),
                               i2 = tri_byte(s_tri[w], 3u * t + 2u);
                const float4 c0 = s_clip[w][i0], c1 = s_clip[w][i1], c2 = s_clip[w][i2];
                if (cached && bary) {
                    // TriCache: r[k] = v_{k+1} x v_{k+2} over the clip (x, y, w) of the
                    // vertices, then the colours; 48 bytes always, the uv words only
                    // for textured frames (written from registers: a local TriCache
                    // would be placed in LDS by the compiler)
                    const float r00 = c1.y * c2.w - c1.w * c2.y, r01 = c1.w * c2.x - c1.x * c2.w,
                                r02 = c1.x * c2.y - c1.y * c2.x;
                    const float r10 = c2.y * c0.w - c2.w * c0.y, r11 = c2.w * c0.x - c2.x * c0.w,
                                r12 = c2.x * c0.y - c2.y * c0.x;
                    const float r20 = c0.y * c1.w - c0.w * c1.y, r21 = c0.w * c1.x - c0.x * c1.w,
                                r22 = c0.x * c1.y - c0.y * c1.x;
                    uint4* dst = reinterpret_cast<uint4*>(a.tri_cache + (size_t)slot * 128u + t);
                    gstore(dst, make_uint4(__float_as_uint(r00), __float_as_uint(r01), __float_as_uint(r02),
                                           __float_as_uint(r10)));
                    gstore(dst + 1, make_uint4(__float_as_uint(r11), __float_as_uint(r12), __float_as_uint(r20),
                                               __float_as_uint(r21)));
                    gstore(dst + 2, make_uint4(__float_as_uint(r22), s_col[w][i0], s_col[w][i1], s_col[w][i2]));
                    if (tex) gstore(dst + 3, make_uint4(s_uv[w][i0], s_uv[w][i1], s_uv[w][i2], 0u));
                }
                static_assert(offsetof(TriCache, col) == 36 && offsetof(TriCache, uv) == 48, "TriCache words");
                live = true;
                if (cull_faces) {
                    const float det = det3_xyw(c0, c1, c2);
                    live = !(tdet < 0.0f ? det < 0.0f : det > 0.0f);
                }
                if (live) {
                    rc.tri();
                    const SVtx v0 = s_sv[w][i0], v1 = s_sv[w][i1], v2 = s_sv[w][i2];
                    if (v0.ok & v1.ok & v2.ok) {
                        const int minx = min(v0.x, min(v1.x, v2.x)), maxx = max(v0.x, max(v1.x, v2.x));
                        const int miny = min(v0.y, min(v1.y, v2.y)), maxy = max(v0.y, max(v1.y, v2.y));
                        live = max((minx - 128 + 255) >> 8, 0) <= min((maxx - 128) >> 8, W - 1)
                               && max((miny - 128 + 255) >> 8, 0) <= min((maxy - 128) >> 8, H - 1);
                    }
                }
                packed = t | (i0 << 8) | (i1 << 16) | (i2 << 24);
            }
            const unsigned long long m = __ballot(live);
            const uint32_t pos = nlive + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (live) s_list[w][pos] = packed;
            nlive += (uint32_t)__popcll(m);
        }
        wave_sync();

        // Stage 2, the packed triangles: clip queue, setup, small or big.
        uint32_t npix[2] = { 0u, 0u };
#pragma unroll
        for (int sl = 0; sl < 2; ++sl) {
            if (64u * (uint32_t)sl >= nlive) break;   // wave-uniform
            const uint32_t pos = lane + 64u * (uint32_t)sl;
            bool want_big = false;
            TriSetup ts;
            uint32_t id = 0;
            do {   // per-lane triangle work; `continue` leaves it
                if (pos >= nlive) continue;
                const uint32_t e = s_list[w][pos];
                const uint32_t t = e & 0xFFu, i0 = (e >> 8) & 0xFFu, i1 = (e >> 16) & 0xFFu, i2 = e >> 24;
                id = (d << VKGV_TRIANGLE_BITS) | t;
                const SVtx v0 = s_sv[w][i0], v1 = s_sv[w][i1], v2 = s_sv[w][i2];
                if (!(v0.ok & v1.ok & v2.ok)) {
                    const uint32_t slot = atomicAdd(&a.ctr->clip_count, 1u);
                    if (slot < a.clip_cap) {
                        // the record's four 16-byte words, from LDS and registers
                        float4* q = reinterpret_cast<float4*>(a.clipq + slot);
                        gstore(q, s_clip[w][i0]);
                        gstore(q + 1, s_clip[w][i1]);
                        gstore(q + 2, s_clip[w][i2]);
                        gstore(reinterpret_cast<uint4*>(q) + 3, make_uint4(id, 0u, 0u, 0u));
                    } else {
                        atomicOr(&a.ctr->overflow, 2u);
                        s_spill[w] = 1u;
                    }
                    continue;
                }
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
                        st.ex[k] = (int16_t)ex;
                        st.ey[k] = (int16_t)ey;
                    }
                    depth_plane(ts.x, ts.y, ts.z, ts.area, st.za, st.zb);
                    st.z0 = ts.z[0];
                    st.ox = (int16_t)(sx0 - ts.x[0]);
                    st.oy = (int16_t)(sy0 - ts.y[0]);
                    st.org = (uint32_t)ts.px0 | ((uint32_t)ts.py0 << 16);
                    const uint32_t bw = (uint32_t)(ts.px1 - ts.px0 + 1), bh = (uint32_t)(ts.py1 - ts.py0 + 1);
                    st.dim = (uint16_t)(bw | (bh << 8));
                    st.pad = 0;
                    st.id = id;
                    s_st[w][pos] = st;
                    npix[sl] = bw * bh;
                } else {
                    want_big = true;
                }
            } while (false);
            if (push_big_wave(want_big, ts, id, lane, bq)) s_spill[w] = 1u;
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

        // A heavy meshlet keeps the first kJobKeep slots and defers the rest
        // to the bin pass in runs of kJobSlots (staged record + jobs), so that
        // its fragment atomics spread over many CUs instead of queueing in
        // this wave's.  Without room for the record or its jobs the wave walks
        // every slot itself.
        uint32_t keep = total;
        if (total > kJobKeep + kJobSlots / 2u) {   // wave-uniform
            const uint32_t njobs = (total - kJobKeep + kJobSlots - 1u) / kJobSlots;
            // one atomic reserves the record (high bits) and the jobs (low bits)
            unsigned long long res = 0ull;
            if (lane == 0) res = atomicAdd(&a.ctr->frag_jobs, (1ull << kJobCountBits) | (unsigned long long)njobs);
            res = __shfl(res, 0);
            const uint32_t sidx = (uint32_t)(res >> kJobCountBits);
            const unsigned long long j0 = res & ((1ull << kJobCountBits) - 1ull);
            const bool room = sidx < a.stage_cap && j0 + njobs <= (unsigned long long)a.job_cap;
            // every reserved job inside the list is written, marked unused without room
#pragma unroll 1
            for (uint32_t q = lane; q < njobs; q += 64u)
                if (j0 + q < (unsigned long long)a.job_cap)
                    gstore(a.jobs + j0 + q, room ? FragJob{ sidx, kJobKeep + q * kJobSlots } : FragJob{ 0xFFFFFFFFu, 0u });
            if (room) {
                uint4* dst = reinterpret_cast<uint4*>(a.stage[sidx].w);
                const uint4* st4 = reinterpret_cast<const uint4*>(s_st[w]);
#pragma unroll 2
                for (uint32_t q = lane; q < kStageTri4; q += 64u) gstore(dst + q, st4[q]);
                if (lane < 32u) gstore(dst + kStageTri4 + lane, reinterpret_cast<const uint4*>(s_pre[w])[lane]);
                if (lane == 0) gstore(dst + kStageTri4 + 32u, make_uint4(total, 0u, 0u, 0u));
                keep = kJobKeep;
            }
        }
        {
            const uint32_t per = (keep + 63u) >> 6;
            const uint32_t k = lane * per;
            walk_small(s_st[w], s_pre[w], total, k, min(k + per, keep), a.keys, a.W, rc);
        }
        wave_sync();
        // the meshlet is redone by the direct path after this kernel
        if (__builtin_amdgcn_readfirstlane(s_spill[w])) {
            if (lane == 0) {
                a.redo[atomicAdd(&a.ctr->redo_count, 1u)] = idx;
                s_spill[w] = 0u;
            }
            wave_sync();
        }
        idx = nidx;      // static stride: no shared cursor
    }
    rc.flush(a.ctr);
}

// ---------------------------------------------------------------------------
struct BigArgs {
    BigTri* big;
    uint32_t* big_chunk;
    uint32_t big_cap;
    uint32_t* chunk_tri;
    uint32_t chunk_cap;
    const ClipTri* clipq;
    uint32_t clip_cap;
    Counters* ctr;
    unsigned long long* keys;
    uint32_t W, H;
    uint32_t tiles_x, tiles_y;
    uint32_t* bin_count;            // this pass's per-tile counts
    uint32_t* bins;
    uint32_t tile_cap;
    BigSetup* setups;
    const FragStage* stage;
    const FragJob* jobs;
    uint32_t job_cap;
};

// One queued triangle per wave at a time: lane 0 clips it in the wave's LDS
// polygon buffers, then lane f - 1 sets up fan triangle (p0, pf, pf+1),
// f = 1 .. m - 2, and the fans join the big triangles with one reservation
// for the wave; a fan without room is swept by the wave.  The block is one
// wave.
template <bool STATS>
__device__ __forceinline__ void clip_view(const BigArgs& a)
{
    RasterCount<STATS> rc;
    __shared__ float4 s_poly[2 * kClipMax];
    const uint32_t lane = threadIdx.x & 63u;
    const LocalPoly pa{ s_poly }, pb{ s_poly + kClipMax };
    const BigQueue q{ a.big, a.big_chunk, a.big_cap, a.chunk_tri, a.chunk_cap, a.ctr, a.keys, a.W };
    const uint32_t n = min(a.ctr->clip_count, a.clip_cap);
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const ClipTri ct = a.clipq[i];
        const uint32_t id = __builtin_amdgcn_readfirstlane(ct.id);
        int m = 0;
        if (lane == 0) {
            m = clip_polygon(pa, pb, ct.c[0], ct.c[1], ct.c[2], a.W, a.H);
            if (m) {
                atomicAdd(&a.ctr->clipped, 1u);
                if (!polygon_projects(pb, m, hw, hh)) m = 0;
            }
        }
        m = __shfl(m, 0);
        wave_sync();   // lane 0's polygon is visible to the wave
        TriSetup t;
        const bool fan = (int)lane + 2 < m && fan_setup(pb, (int)lane + 1, hw, hh, (int)a.W, (int)a.H, t);
        const bool spill = push_big_wave(fan, t, id, lane, q);
        if (__ballot(spill) != 0ull) wave_sweep(spill, t, id, a.keys, a.W, rc);
        wave_sync();   // the polygon buffers are reused by the wave's next triangle
    }
    rc.flush(a.ctr);
}

__device__ __forceinline__ uint32_t chunks_of(const BigTri& t)
{
    const int px0 = (int)(t.bb0 & 0xFFFFu), py0 = (int)(t.bb0 >> 16);
    const int px1 = (int)(t.bb1 & 0xFFFFu), py1 = (int)(t.bb1 >> 16);
    const uint32_t ntx = (uint32_t)(px1 / (int)kBinTile - px0 / (int)kBinTile + 1);
    const uint32_t nty = (uint32_t)(py1 / (int)kBinTile - py0 / (int)kBinTile + 1);
    return (ntx * nty + 63u) >> 6;
}

// Big-triangle slots of this pass (stored or not) and the chunks of the chunk
// list in use.
__device__ __forceinline__ void pass_extent(const BigArgs& a, uint32_t& nbig, uint32_t& nchunks)
{
    nbig = min(a.ctr->big_count, a.big_cap);
    nchunks = (uint32_t)min(a.ctr->big_chunks, (unsigned long long)a.chunk_cap);
}

// A deferred fragment run (one wave): lane l walks its share of the job's
// kJobSlots slots, reading the staged triangles and slot prefix straight from
// the record in global memory.  (Staging them through LDS measured slower:
// 30 KB of static LDS made every launch of this 1024-block kernel about
// 10 us longer, empty ones included.)
template <bool STATS>
__device__ __forceinline__ void frag_job(const BigArgs& a, uint32_t j, RasterCount<STATS>& rc)
{
    const uint32_t lane = threadIdx.x & 63u;
    const FragJob job = a.jobs[j];
    const uint32_t stage = __builtin_amdgcn_readfirstlane(job.stage);
    const uint32_t first = __builtin_amdgcn_readfirstlane(job.first);
    if (stage == 0xFFFFFFFFu) return;   // reserved by a meshlet that found no room
    const uint32_t* rec = a.stage[stage].w;
    const SmallTri* tri = reinterpret_cast<const SmallTri*>(rec);
    const uint32_t* pre = rec + kStageTriWords;
    const uint32_t total = __builtin_amdgcn_readfirstlane(rec[kStageTriWords + 128u]);
    const uint32_t end = min(first + kJobSlots, total);
    const uint32_t per = (end - first + 63u) >> 6;
    const uint32_t k = first + lane * per;
    walk_small(tri, pre, total, k, min(k + per, end), a.keys, a.W, rc);
}

// The bin pass's work items: the 64-tile chunks of the big triangles, then
// the deferred fragment runs of heavy meshlets.  Waves are numbered
// block-minor, so consecutive items (chunks of one triangle, runs of one
// meshlet) run on different CUs.
//
// One wave per 64-tile chunk of a big triangle's bounding-box tiles.  The
// wave of chunk 0 stores the triangle's rasterization setup.  Lane l tests
// tile 64*chunk + l: a tile is skipped when one edge is negative at all four
// corner pixel centres of the tile's part of the bounding box (no centre there
// can be covered); a kept tile appends the triangle to its list.
//
// The wave takes items first, first + nw, ... of the view's nitems (chunks,
// then runs).
template <bool STATS>
__device__ __forceinline__ void bin_view(const BigArgs& a, uint32_t first, uint32_t nw, uint32_t nitems)
{
    RasterCount<STATS> rc;
    uint32_t pairs = 0u, most = 0u;   // (tile, triangle) pairs binned, longest tile list (STATS)
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t nbig, nchunks;
    pass_extent(a, nbig, nchunks);
    for (uint32_t c = first; c < nitems; c += nw) {
        if (c >= nchunks) {
            frag_job<STATS>(a, c - nchunks, rc);
            continue;
        }
        // the chunk is in use iff its entry names a slot stored in this pass
        // whose range holds it (chunks reserved by a triangle that found no
        // room keep an entry of an earlier pass, or none); uniform over the wave
        const uint32_t ti = a.chunk_tri[c];
        const bool known = ti < nbig;
        uint32_t c0 = 0xFFFFFFFFu;
        BigTri t;
        if (known) {   // both loads in flight together
            c0 = a.big_chunk[ti];
            t = a.big[ti];
        }
        const bool used = known && c0 <= c && c - c0 < chunks_of(t);
        if (!used) continue;
        const uint32_t j = c - c0;
        Edge e[3];
        edges_of(t.x, t.y, e);
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
        bool keep = false;
        uint32_t tile = 0u;
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
            keep = !out;
            tile = (uint32_t)ty * a.tiles_x + (uint32_t)tx;
        }
        // the triangle joins its tile's list, or, when the list is full, is
        // swept over that tile right here (the tile pass, a later launch,
        // merges its own list over these keys)
        uint32_t pos = 0u;
        if (keep) {
            pos = atomicAdd(&a.bin_count[tile], 1u);
            if (pos < a.tile_cap) a.bins[(size_t)tile * a.tile_cap + pos] = ti;
        }
        const bool spill = keep && pos >= a.tile_cap;
        if constexpr (STATS) {
            pairs += keep ? 1u : 0u;
            most = max(most, keep ? pos + 1u : 0u);
        }
        if (__ballot(spill) != 0ull) {   // wave-uniform
            if (lane == 0) atomicOr(&a.ctr->overflow, 8u);
            const long long area = ((long long)t.x[1] - t.x[0]) * ((long long)t.y[2] - t.y[0])
                                   - ((long long)t.y[1] - t.y[0]) * ((long long)t.x[2] - t.x[0]);
            float za, zb;
            depth_plane(t.x, t.y, t.z, area, za, zb);
            unsigned long long pending = __ballot(spill);
            while (pending != 0ull) {
                const int sl = __ffsll((long long)pending) - 1;
                pending &= pending - 1ull;
                const uint32_t st = __shfl(tile, sl);
                const int rx = (int)(st % a.tiles_x) * (int)kBinTile, ry = (int)(st / a.tiles_x) * (int)kBinTile;
                sweep_rect(e, t.x, t.y, t.z, za, zb, t.id, max(rx, px0), max(ry, py0),
                           min(rx + (int)kBinTile - 1, px1), min(ry + (int)kBinTile - 1, py1), a.keys, a.W, rc);
            }
        }
    }
    if constexpr (STATS) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            pairs += __shfl_xor(pairs, off);
            most = max(most, __shfl_xor(most, off));
        }
        if (lane == 0 && pairs) {
            atomicAdd(&a.ctr->big_total_pairs, pairs);
            atomicMax(&a.ctr->max_tile_tris, most);
        }
    }
    rc.flush(a.ctr);
}

// Tile pass.  Work item (tile, band, part): band b (32x16 pixels) of a
// screen tile, rasterizing part p of its list of n triangles
// (triangles [n p / S, n (p + 1) / S), S = min(kTileSplit, ceil(n / item
// tris)) parts).  A band keeps its keys in registers (lane l owns column
// l & 31, rows (l >> 5) + 2r, r = 0..7 of the band), then merges them into
// the visibility buffer: plain read-min-store when the tile has one part (the
// meshlet raster and the bin pass of this pass have finished and bands are
// disjoint; the old keys are read when the item starts), atomicMin
// otherwise.  Setups are loaded 64 at a time, one per lane, and broadcast
// with readlane.
//
// The items are enumerated statically, part-major: item i is part
// i / (2 tiles) of band i % (2 tiles); wave w takes items w, w + waves, ...,
// 64 at a time: lane l reads the count of item r + l's tile and the wave runs
// the items whose part the count has.  The items of one tile are 2 tiles
// apart, so they run in different waves.
constexpr int kTileParts = 2;
constexpr int kTileRows = (int)(kBinTile * kBinTile / 64u) / kTileParts;     // pixels per lane in a band
constexpr int kSetupWords = (int)(sizeof(BigSetup) / 4);
static_assert(kSetupWords % 4 == 0, "setups are loaded 16 bytes at a time");

// Parts of a tile whose list holds n triangles (0 for an empty tile).
__device__ __forceinline__ uint32_t tile_split(uint32_t n, uint32_t item_tris)
{
    return min(kTileSplit, (n + item_tris - 1u) / item_tris);
}

// One work item: band `band` of screen tile `tile`, triangles [b, e) of the
// tile's list (wave-uniform arguments).
template <bool STATS>
__device__ __forceinline__ void tile_item(const BigArgs& a, uint32_t tile, uint32_t band, uint32_t b, uint32_t e,
                                          uint32_t split, RasterCount<STATS>& rc)
{
    const uint32_t lane = threadIdx.x & 63u;
    const int lx = (int)(lane & 31u), ly = (int)(lane >> 5);
    const BigSetup* __restrict__ setups = a.setups;
    const int px = (int)((tile % a.tiles_x) * kBinTile) + lx;
    const int py0 = (int)((tile / a.tiles_x) * kBinTile + band * (kBinTile / kTileParts)) + ly;
    const bool col_in = px < (int)a.W;
    unsigned long long old[kTileRows];
    if (split == 1) {
#pragma unroll
        for (int r = 0; r < kTileRows; ++r) {
            const int py = py0 + 2 * r;
            old[r] = col_in && py < (int)a.H ? a.keys[(size_t)py * a.W + px] : 0ull;
        }
    }
    unsigned long long key[kTileRows];
#pragma unroll
    for (int r = 0; r < kTileRows; ++r) key[r] = VKGV_KEY_CLEAR;
    const int sx = px * 256 + 128;
    for (uint32_t t0 = b; t0 < e; t0 += 64u) {
        // batch of up to 64 triangles: lane j holds the setup of triangle t0 + j
        // (two loads for the batch), broadcast one at a time below
        const uint32_t m = min(64u, e - t0);
        uint32_t sw[kSetupWords];
        if (lane < m) {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(setups + a.bins[t0 + lane]);
#pragma unroll
            for (int q = 0; q < kSetupWords; q += 4) {
                const uint4 v4 = *reinterpret_cast<const uint4*>(src + q);
                sw[q] = v4.x; sw[q + 1] = v4.y; sw[q + 2] = v4.z; sw[q + 3] = v4.w;
            }
        }
        for (uint32_t j = 0; j < m; ++j) {
            BigSetup s;
            uint32_t* dst = reinterpret_cast<uint32_t*>(&s);
#pragma unroll
            for (int q = 0; q < kSetupWords; ++q) dst[q] = __builtin_amdgcn_readlane(sw[q], j);
            const int bx0 = (int)(s.bb0 & 0xFFFFu), by0 = (int)(s.bb0 >> 16);
            const int bx1 = (int)(s.bb1 & 0xFFFFu);
            const bool col_ok = px >= bx0 && px <= bx1;
            long long ev[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                ev[q] = s.e0[q] - (long long)(px - bx0) * (long long)s.ey[q] * 256
                        + (long long)(py0 - by0) * (long long)s.ex[q] * 256;
            const float dxf = (float)(sx - s.x0) * 0.00390625f;
            if (!col_ok) continue;
            // Rows without branches: a centre is covered iff its three edge
            // functions are >= 0 (the sign of their OR's high word).  A covered
            // centre lies in the triangle's bounding box, so no box test is
            // needed; centres outside the view are never stored.  Depth is
            // the same expression as everywhere, with za * dxf formed once.
            const float zx = s.za * dxf;
            const int ybase = py0 * 256 + 128 - s.y0;
            const uint32_t idw = 0xFFFFFFFEu - s.id;
#pragma unroll
            for (int r = 0; r < kTileRows; ++r) {
                if (r) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) ev[q] += (long long)s.ex[q] * 512;   // +2 rows
                }
                const int hi = (int)(ev[0] >> 32) | (int)(ev[1] >> 32) | (int)(ev[2] >> 32);
                const float dyf = (float)(ybase + 512 * r) * 0.00390625f;
                const float z = s.z0 + (zx + s.zb * dyf);
                const bool ok = (hi >= 0) & (z >= 0.0f) & (z <= 1.0f);
                // an uncovered centre's key is the clear key, never below key[r]
                const uint32_t khi = ok ? ~__float_as_uint(z + 0.0f) : 0xFFFFFFFFu;
                const uint32_t klo = ok ? idw : 0xFFFFFFFFu;
                const unsigned long long k = ((unsigned long long)khi << 32) | klo;
                key[r] = k < key[r] ? k : key[r];
                if constexpr (STATS) {
                    if (ok && py0 + 2 * r < (int)a.H) rc.frag();   // centres below the view are never stored
                }
            }
        }
    }
    if (!col_in) return;
#pragma unroll
    for (int r = 0; r < kTileRows; ++r) {
        const int py = py0 + 2 * r;
        if (key[r] == VKGV_KEY_CLEAR || py >= (int)a.H) continue;
        unsigned long long* p = a.keys + (size_t)py * a.W + px;
        if (split == 1) {
            if (key[r] < old[r]) *p = key[r];
            rc.store();
        } else {
            key_min(p, key[r]);
            rc.atomic();
        }
    }
}

// Views of more than kScanTiles tiles: the items are enumerated statically,
// part-major: item i is part i / (2 tiles) of band i % (2 tiles); wave w takes
// items w, w + waves, ..., 64 at a time: lane l reads the count of item r + l's
// tile and the wave runs the items whose part the count has.
template <bool STATS>
__device__ __forceinline__ void tile_view_static(const BigArgs& a, uint32_t item_tris, RasterCount<STATS>& rc)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t wave = blockIdx.x * wpb + (threadIdx.x >> 6), nwaves = gridDim.x * wpb;
    const uint32_t nband = a.tiles_x * a.tiles_y * (uint32_t)kTileParts;
    const unsigned long long nitems = (unsigned long long)nband * kTileSplit;
    for (unsigned long long r0 = 0; wave + r0 * nwaves < nitems; r0 += 64u) {
        const unsigned long long it = wave + (r0 + lane) * nwaves;
        uint32_t code = 0u, n = 0u;
        bool live = false;
        if (it < nitems) {
            const uint32_t part = (uint32_t)(it / nband), hv = (uint32_t)(it - (unsigned long long)part * nband);
            n = min(a.bin_count[hv / (uint32_t)kTileParts], a.tile_cap);
            live = part < tile_split(n, item_tris);
            code = hv | (part << 28);
        }
        unsigned long long pending = __ballot(live);
        while (pending != 0ull) {
            const int j = __ffsll((long long)pending) - 1;
            pending &= pending - 1ull;
            const uint32_t cj = __builtin_amdgcn_readlane(code, j), nj = __builtin_amdgcn_readlane(n, j);
            const uint32_t hv = cj & 0x0FFFFFFFu, part = cj >> 28;
            const uint32_t tile = hv / (uint32_t)kTileParts, split = tile_split(nj, item_tris);
            const uint32_t start = tile * a.tile_cap;
            tile_item<STATS>(a, tile, hv % (uint32_t)kTileParts, start + nj * part / split,
                             start + nj * (part + 1u) / split, split, rc);
        }
    }
}

// Views of at most kScanTiles tiles (4K included): every block scans the
// tile counts itself (kScanPer tiles per thread, in registers, 16-byte
// loads), so the items form one compact list, and wave g of the grid takes
// (item, band) pairs g, g + waves, ...  When the list has at most kScanTiles
// items and no list is longer than kItemMaxN, the scan writes each item's
// record (tile | part << 13 | (split - 1) << 16 | n << 19) to LDS, one read
// per item; otherwise it writes the tiles' first items and a wave finds an
// item's tile by binary search.  The redundant per-block scan costs a few
// microseconds; it replaces a launch, and the compact list keeps every wave
// within one item of the average.  In a batch, a view's items continue the
// previous scanned view's wave rotation, so views with few items each still
// spread over the grid.
constexpr uint32_t kTileThreads = 1024;
constexpr uint32_t kScanTiles = 8192;
constexpr uint32_t kScanPer = kScanTiles / kTileThreads;   // tiles per thread in the scan
constexpr uint32_t kItemMaxN = 2047;                       // longest list an LDS item record holds
static_assert(kScanPer % 4 == 0, "the scan loads whole uint4 runs per thread");
static_assert(kScanTiles <= (1u << 13) && kTileSplit <= 8, "item records: 13-bit tile, 3-bit part and split");

template <bool STATS>
__device__ __forceinline__ void tile_view_scan(const BigArgs& a, uint32_t item_tris, uint32_t* s_arr,
                                               uint32_t* s_wsum, uint32_t& carry, RasterCount<STATS>& rc)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t ntiles = a.tiles_x * a.tiles_y;   // at most kScanTiles
    const uint32_t b0 = tid * kScanPer;
    uint32_t cnt[kScanPer];
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; q += 4) {
        if (b0 + q + 4 <= ntiles) {
            const uint4 v4 = *reinterpret_cast<const uint4*>(a.bin_count + b0 + q);
            cnt[q] = v4.x; cnt[q + 1] = v4.y; cnt[q + 2] = v4.z; cnt[q + 3] = v4.w;
        } else {
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) cnt[q + r] = b0 + q + r < ntiles ? a.bin_count[b0 + q + r] : 0u;
        }
    }
    uint32_t isum = 0;
#pragma unroll
    for (uint32_t u = 0; u < kScanPer; ++u) {
        cnt[u] = min(cnt[u], a.tile_cap);
        isum += tile_split(cnt[u], item_tris);
    }
    uint32_t x = isum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) s_wsum[w] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (uint32_t q = 0; q < kTileThreads / 64; ++q) {
        if (q < w) wbase += s_wsum[q];
        total += s_wsum[q];
    }
    // records per item, or first item per tile (block-uniform choice)
    const bool records = total <= kScanTiles && a.tile_cap <= kItemMaxN;
    uint32_t run = wbase + x - isum;
#pragma unroll
    for (uint32_t u = 0; u < kScanPer; ++u) {
        const uint32_t sp = tile_split(cnt[u], item_tris);
        if (records) {
            for (uint32_t p = 0; p < sp; ++p)
                s_arr[run + p] = (b0 + u) | (p << 13) | ((sp - 1u) << 16) | (cnt[u] << 19);
        } else {
            s_arr[b0 + u] = run;
        }
        run += sp;
    }
    __syncthreads();
    const uint32_t nwaves = gridDim.x * (kTileThreads / 64);
    // the view's items start where the previous view's ended (mod waves)
    const uint32_t first = (blockIdx.x + gridDim.x * w + nwaves - carry) % nwaves;
    carry = (carry + kTileParts * total) % nwaves;
    for (uint32_t hv = first; hv < kTileParts * total; hv += nwaves) {
        const uint32_t item = hv / (uint32_t)kTileParts;
        uint32_t tile, part, split, n;
        if (records) {
            const uint32_t r = __builtin_amdgcn_readfirstlane(s_arr[item]);
            tile = r & 0x1FFFu;
            part = (r >> 13) & 7u;
            split = ((r >> 16) & 7u) + 1u;
            n = r >> 19;
        } else {
            uint32_t lo = 0, hi = ntiles - 1u;   // last tile whose first item is <= item
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1u) >> 1;
                if (s_arr[mid] <= item) lo = mid; else hi = mid - 1u;
            }
            tile = __builtin_amdgcn_readfirstlane(lo);
            part = item - s_arr[tile];
            n = __builtin_amdgcn_readfirstlane(min(a.bin_count[tile], a.tile_cap));
            split = tile_split(n, item_tris);
        }
        const uint32_t start = tile * a.tile_cap;
        tile_item<STATS>(a, tile, hv % (uint32_t)kTileParts, start + n * part / split, start + n * (part + 1u) / split,
                         split, rc);
    }
    __syncthreads();   // the LDS array is reused by the next view
}

// ---------------------------------------------------------------------------
// Kernels over a batch of views: every view's work in one launch.

// The view's big-triangle path in raster pass `pass` (0: first, 1: late).
__device__ __forceinline__ BigArgs big_args(const ViewDev& v, uint32_t pass)
{
    BigArgs a;
    a.big = v.big;
    a.big_chunk = v.big_chunk;
    a.big_cap = v.big_capacity;
    a.chunk_tri = v.chunk_tri;
    a.chunk_cap = v.chunk_capacity;
    a.clipq = v.clipq;
    a.clip_cap = v.clip_capacity;
    a.ctr = v.counters;
    a.keys = v.keys;
    a.W = v.width;
    a.H = v.height;
    a.tiles_x = v.tiles_x;
    a.tiles_y = v.tiles_y;
    a.bin_count = v.bin_count + (size_t)pass * tile_count_stride(v);
    a.bins = v.bins;
    a.tile_cap = v.tile_capacity;
    a.setups = v.setups;
    a.stage = v.frag_stage;
    a.jobs = v.frag_jobs;
    a.job_cap = v.job_capacity;
    return a;
}

struct RasterBatchArgs {
    SceneDev s;
    BatchDev b;
    uint32_t flags;
    uint32_t list;
    int small_lim;
    uint32_t cache_slots;           // tuning: meshlets per view cached for the resolve
};

// The views of a batch are visited 64 at a time: lane l of each wave reads
// the counter of view v0 + l, so that a view without work for the wave costs
// one readlane instead of a dependent load.  The views' lists are dealt out
// as one concatenated list: view j's meshlet i goes to wave
// (i + meshlets of the views before j) mod waves, so every wave gets within
// one meshlet of the average however the meshlets spread over the views.
// Waves are numbered block-minor (wave w of block b is wave w * blocks + b),
// so consecutive meshlets of a list, which are neighbours on screen and
// similar in cost, run in different blocks and thus on different CUs and
// XCDs: the fragment atomics of a cluster of heavy meshlets near the camera
// do not queue in one CU's memory pipeline.
// Three waves per SIMD: the register budget (168) the kernel fits without spilling.
template <bool STATS>
__global__ __launch_bounds__(kBlock, 3) void raster_kernel(RasterBatchArgs ab)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = (threadIdx.x >> 6) * gridDim.x + blockIdx.x;
    const uint32_t nwaves = gridDim.x * kWaves;
    uint32_t carry = 0;   // meshlets of the views before v0, mod nwaves
    for (uint32_t v0 = 0; v0 < ab.b.nviews; v0 += 64u) {
        uint32_t cnt = 0, base = 0;
        if (v0 + lane < ab.b.nviews) {
            const Counters* c = batch_view(ab.b, v0 + lane).counters;
            cnt = c->vis_count[ab.list];
            base = ab.list ? c->vis_count[0] : 0u;
        }
        // exclusive prefix of the counts over the 64 views, modulo nwaves
        uint32_t incl = cnt % nwaves;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl = (incl + y) % nwaves;
        }
        const uint32_t excl = (incl + nwaves - cnt % nwaves) % nwaves;
        const uint32_t m = min(64u, ab.b.nviews - v0);
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t count = __builtin_amdgcn_readlane(cnt, j);
            const uint32_t rot = (carry + __builtin_amdgcn_readlane(excl, j)) % nwaves;
            const uint32_t first = (gw + nwaves - rot) % nwaves;
            if (first >= count) continue;
            const uint32_t vi = v0 + j;
            const ViewDev& v = batch_view(ab.b, vi);
            // the view's fields are the same for the whole wave: held in
            // scalar registers (the compiler cannot tell from the vector
            // loads that read them)
            RasterArgs a;
            a.s = ab.s;
            a.list = uni(v.visible[ab.list]);
            a.keys = uni(v.keys);
            a.W = uni(v.width);
            a.H = uni(v.height);
            a.flags = ab.flags;
            a.small_lim = ab.small_lim;
            a.big = uni(v.big);
            a.big_chunk = uni(v.big_chunk);
            a.big_cap = uni(v.big_capacity);
            a.redo = uni(v.redo);
            a.chunk_tri = uni(v.chunk_tri);
            a.chunk_cap = uni(v.chunk_capacity);
            a.clipq = uni(v.clipq);
            a.clip_cap = uni(v.clip_capacity);
            a.ctr = uni(v.counters);
            a.mvp = uni(v.mvp);
            a.count = count;
            a.first = uni(first);
            a.slot0 = __builtin_amdgcn_readlane(base, j);
            a.tri_cache = uni(v.tri_cache);
            a.slot_albedo = uni(v.slot_albedo);
            a.slot_mat = uni(v.slot_mat);
            a.draw_slot = uni(v.draw_slot);
            a.cache_slots = uni(min(v.cache_slots, ab.cache_slots));
            a.stage = uni(v.frag_stage);
            a.jobs = uni(v.frag_jobs);
            a.stage_cap = uni(v.stage_capacity);
            a.job_cap = uni(v.job_capacity);
            raster_view<STATS>(a);
        }
        carry = (carry + __builtin_amdgcn_readlane(incl, 63)) % nwaves;
    }
}

// Lane l of the wave reads field(view v0 + l) (0 past the batch).
template <typename F>
__device__ __forceinline__ uint32_t lane_view_value(const BatchDev& b, uint32_t v0, F field)
{
    const uint32_t v = v0 + (threadIdx.x & 63u);
    return v < b.nviews ? field(*batch_view(b, v).counters) : 0u;
}

__device__ __forceinline__ uint32_t big_count(const Counters& c) { return c.big_count; }

// What the direct path needs to redo a pass's meshlets.
struct RedoArgs {
    SceneDev s;
    uint32_t flags;
    uint32_t list;
    int small_lim;
};

// The meshlets of this pass that had a triangle without room in a queue: one
// wave per meshlet, one lane per triangle, with the raster's own vertex,
// culling and setup arithmetic; every triangle of theirs that the raster did
// not keep in its wave (not small) goes to the wave's direct sweep, clipped
// first (in the lane's private memory) when it needs clipping.  A triangle
// that did reach a queue is written twice; the visibility write is a
// minimum, so the result is the same.  Every lane runs the same iterations.
template <bool STATS>
__device__ __forceinline__ void redo_view(const RedoArgs& ra, const ViewDev& v, uint32_t n)
{
    RasterCount<STATS> rc;
    const RasterRec* list = v.visible[ra.list];
    const float hw = (float)v.width * 0.5f, hh = (float)v.height * 0.5f;
    const int W = (int)v.width, H = (int)v.height;
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        const RasterRec rec = list[v.redo[r]];
        const float* mrow = v.mvp + (size_t)rec.transform * kMvpStride;
        const Mat4 mvp = load_mat4(mrow);
        const float tdet = mrow[16];
        const bool cull_faces = !(rec.material >> 31) && !(ra.flags & VKGV_NO_BACKFACE);
        const uint32_t tc = (rec.counts >> 8) & 0xFFu;
        const uint32_t* words = reinterpret_cast<const uint32_t*>(ra.s.tri_base + rec.tri);
        const float* vtx = reinterpret_cast<const float*>(reinterpret_cast<const uint8_t*>(ra.s.vtx_base) + rec.vtx);
        for (uint32_t t0 = 0; t0 < tc; t0 += 64u) {
            const uint32_t t = t0 + lane;
            const uint32_t id = (rec.draw << VKGV_TRIANGLE_BITS) | t;
            float4 c[3];
            bool live = t < tc, clip = false, direct = false;
            TriSetup ts;
            if (live) {
                SVtx sv[3];
                bool ok = true;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint32_t vi = ra.s.vidx_base[rec.vidx + tri_byte(words, 3u * t + (uint32_t)k)];
                    const float* pp = vtx + (size_t)vi * 6;
                    c[k] = mat4_point(mvp, pp[0], pp[1], pp[2]);
                    sv[k].x = 0; sv[k].y = 0; sv[k].z = 0.0f;
                    sv[k].ok = project_vtx(c[k], hw, hh, sv[k]) ? 1u : 0u;
                    ok = ok && sv[k].ok;
                }
                if (cull_faces) {
                    const float det = det3_xyw(c[0], c[1], c[2]);
                    if (tdet < 0.0f ? det < 0.0f : det > 0.0f) live = false;
                }
                if (live && !ok) {
                    clip = true;
                } else if (live && setup_tri(sv[0], sv[1], sv[2], W, H, ts)) {
                    direct = !(ts.maxx - ts.minx < ra.small_lim && ts.maxy - ts.miny < ra.small_lim);
                }
            }
            wave_sweep(direct, ts, id, v.keys, v.width, rc);
            if (__ballot(clip) == 0ull) continue;   // wave-uniform
            float4 A[kClipMax], B[kClipMax];
            const LocalPoly pa{ A }, pb{ B };
            int m = 0;
            if (clip) {
                m = clip_polygon(pa, pb, c[0], c[1], c[2], v.width, v.height);
                if (m && !polygon_projects(pb, m, hw, hh)) m = 0;
            }
            for (int f = 1; f + 1 < kClipMax; ++f) {
                if (__ballot(f + 1 < m) == 0ull) break;   // wave-uniform
                TriSetup ft;
                const bool want = f + 1 < m && fan_setup(pb, f, hw, hh, W, H, ft);
                wave_sweep(want, ft, id, v.keys, v.width, rc);
            }
        }
    }
    rc.flush(v.counters);
}

// Per view: the clip queue, then the meshlets to redo.
template <bool STATS>
__global__ __launch_bounds__(kClipThreads) void clip_kernel(BatchDev b, RedoArgs ra)
{
    for (uint32_t v0 = 0; v0 < b.nviews; v0 += 64u) {
        const uint32_t n = lane_view_value(b, v0, [](const Counters& c) { return c.clip_count; });
        const uint32_t nr = lane_view_value(b, v0, [](const Counters& c) { return c.redo_count; });
        const uint32_t m = min(64u, b.nviews - v0);
        for (uint32_t j = 0; j < m; ++j) {
            if (__builtin_amdgcn_readlane(n, j)) clip_view<STATS>(big_args(batch_view(b, v0 + j), ra.list));
            const uint32_t r = __builtin_amdgcn_readlane(nr, j);
            if (r) redo_view<STATS>(ra, batch_view(b, v0 + j), r);
        }
    }
}

// Per view: the big triangles' chunks, then the deferred fragment runs.  The
// views' items are dealt out as one concatenated list, as the raster kernel
// deals its meshlets (view j's item i goes to wave (i + items of the views
// before j) mod waves), so that a batch of views with few items each keeps
// the whole grid busy instead of running the views one after another.
template <bool STATS>
__global__ __launch_bounds__(256) void bin_kernel(BatchDev b, uint32_t pass)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = (threadIdx.x >> 6) * gridDim.x + blockIdx.x, nw = gridDim.x * 4u;
    uint32_t carry = 0;   // items of the views before v0, mod nw
    for (uint32_t v0 = 0; v0 < b.nviews; v0 += 64u) {
        uint32_t cnt = 0;
        if (v0 + lane < b.nviews) {
            const ViewDev& v = batch_view(b, v0 + lane);
            const Counters& c = *v.counters;
            const uint32_t nbig = min(c.big_count, v.big_capacity);
            const uint32_t nchunks = (uint32_t)min(c.big_chunks, (unsigned long long)v.chunk_capacity);
            const uint32_t njobs =
                (uint32_t)min(c.frag_jobs & ((1ull << kJobCountBits) - 1ull), (unsigned long long)v.job_capacity);
            cnt = nchunks + njobs;
            if (gw == 0 && nbig) atomicAdd(&v.counters->big_total_tris, nbig);
        }
        uint32_t incl = cnt % nw;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off);
            if (lane >= off) incl = (incl + y) % nw;
        }
        const uint32_t excl = (incl + nw - cnt % nw) % nw;
        const uint32_t m = min(64u, b.nviews - v0);
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t count = __builtin_amdgcn_readlane(cnt, j);
            const uint32_t rot = (carry + __builtin_amdgcn_readlane(excl, j)) % nw;
            const uint32_t first = (gw + nw - rot) % nw;
            if (first < count) bin_view<STATS>(big_args(batch_view(b, v0 + j), pass), first, nw, count);
        }
        carry = (carry + __builtin_amdgcn_readlane(incl, 63)) % nw;
    }
}

// Per view with big triangles in this pass: its tile items.
template <bool STATS>
__global__ __launch_bounds__(kTileThreads) void tile_kernel(BatchDev b, uint32_t pass, uint32_t item_tris)
{
    __shared__ uint32_t s_arr[kScanTiles];
    __shared__ uint32_t s_wsum[kTileThreads / 64];
    RasterCount<STATS> rc;
    uint32_t carry = 0;   // items of the scanned views before this one, mod waves
    for (uint32_t v0 = 0; v0 < b.nviews; v0 += 64u) {
        const uint32_t n = lane_view_value(b, v0, big_count);
        const uint32_t m = min(64u, b.nviews - v0);
        for (uint32_t j = 0; j < m; ++j) {
            if (!__builtin_amdgcn_readlane(n, j)) continue;
            const ViewDev& v = batch_view(b, v0 + j);
            const BigArgs a = big_args(v, pass);
            if (v.tiles_x * v.tiles_y <= kScanTiles) tile_view_scan<STATS>(a, item_tris, s_arr, s_wsum, carry, rc);
            else tile_view_static<STATS>(a, item_tris, rc);
            rc.flush(a.ctr);
        }
    }
}

} // namespace

void launch_raster(const SceneDev& s, const BatchDev& b, uint32_t flags, uint32_t list, uint32_t grid_blocks,
                   uint32_t small_max_extent, uint32_t cache_slots, hipStream_t st)
{
    RasterBatchArgs a;
    a.s = s;
    a.b = b;
    a.flags = flags;
    a.list = list;
    a.small_lim = (int)(small_max_extent * 256u);
    a.cache_slots = cache_slots;
    if (flags & VKGV_STATS) raster_kernel<true><<<grid_blocks, kBlock, 0, st>>>(a);
    else raster_kernel<false><<<grid_blocks, kBlock, 0, st>>>(a);
}

void launch_raster_big(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, uint32_t list,
                       uint32_t small_max_extent, uint32_t tile_grid, uint32_t tile_item_tris, hipStream_t st)
{
    (void)g;
    RedoArgs ra;
    ra.s = s;
    ra.flags = flags;
    ra.list = list;
    ra.small_lim = (int)(small_max_extent * 256u);
    // one wave per block: enough waves for the redo path, which is rare
    const bool stats = flags & VKGV_STATS;
    if (stats) clip_kernel<true><<<std::max(64u, tile_grid / 4u), kClipThreads, 0, st>>>(b, ra);
    else clip_kernel<false><<<std::max(64u, tile_grid / 4u), kClipThreads, 0, st>>>(b, ra);
    if (stats) bin_kernel<true><<<1024, 256, 0, st>>>(b, list);
    else bin_kernel<false><<<1024, 256, 0, st>>>(b, list);
    // one round of the tile kernel's 1024-thread blocks at its occupancy
    const uint32_t tile_blocks = std::max(1u, tile_grid / 8u);
    if (stats) tile_kernel<true><<<tile_blocks, kTileThreads, 0, st>>>(b, list, tile_item_tris);
    else tile_kernel<false><<<tile_blocks, kTileThreads, 0, st>>>(b, list, tile_item_tris);
}

} // namespace vkgv
