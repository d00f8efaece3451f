// Meshlet cull + compaction (visbuffer.task.glsl:25-76, culling.h.glsl:8-56).
//
// A persistent grid of waves strides over work units of 64 * P consecutive
// draws and a group of views.  A lane reads the 32-byte cull records of its
// P draws (each load of the wave one contiguous 2 KiB span), transforms the
// meshlet AABBs to world space once, then for every view of the group runs
// the frustum test, the optional normal-cone test and the Hi-Z probe, and
// appends the unit's visible draws to the view's list with one atomicAdd per
// (unit, view) (ballot + mbcnt prefix), issued for all the group's views
// before any is waited on.  The list order across units is unspecified; the
// visible SET is what the reference defines (its per-task-workgroup payload is
// re-packed by the mesh-shader launch anyway).
#include <algorithm>

#include "internal.hpp"

namespace vkgv {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kReserveChunks = 4;   // chunks per list reservation

// aabbPositions (culling.h.glsl:32-41).
__constant__ float kCorner[8][3] = {
    { 1, -1, -1 }, { 1, 1, -1 }, { -1, 1, -1 }, { -1, -1, -1 },
    { 1, -1, 1 },  { 1, 1, 1 },  { -1, -1, 1 }, { -1, 1, 1 },
};

__device__ __forceinline__ bool aabb_in_frustum(const float c[3], const float e[3], const vkgv_camera& cam)
{
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float* p = cam.frustum[i];
        const float radius = (e[0] * fabsf(p[0]) + e[1] * fabsf(p[1])) + e[2] * fabsf(p[2]);
        const float distance = ((p[0] * c[0] + p[1] * c[1]) + p[2] * c[2]) - p[3];
        if (-radius > distance) return false;
    }
    return true;
}

// A Hi-Z probe, ready to fetch: the 2x2 footprint of the AABB's screen
// centre in the NEAREST mip of its extent, and the AABB's nearest depth.
struct ProbeAt {
    uint32_t off;       // first texel of the level
    int lw;             // level width
    int x0, x1, y0, y1; // footprint texels
    float maxz;
};

// projectAabb (culling.h.glsl:44-56), the probe's level and coordinates
// (visbuffer.task.glsl:56-59), and textureLod's mip selection through the MIN
// sampler (application.cpp:437-458): lod clamped to [0, 16] (NaN -> 0),
// NEAREST mip ceil(lod + 0.5) - 1, clamped to the last level.  The level
// sizes and offsets are read from lanes: lane k holds level k's (pw, ph, po);
// every lane of the wave calls this.
__device__ __forceinline__ ProbeAt probe_setup(const float c[3], const float e[3], const Mat4& vp, float w0, float h0,
                                               int pw, int ph, uint32_t po, uint32_t last)
{
    float mnx = 1.0f, mny = 1.0f, mnz = 1.0f, mxx = -1.0f, mxy = -1.0f, mxz = -1.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float px = kCorner[i][0] * e[0] + c[0];
        const float py = kCorner[i][1] * e[1] + c[1];
        const float pz = kCorner[i][2] * e[2] + c[2];
        const float4 clip = mat4_point(vp, px, py, pz);
        const float nx = gclamp(clip.x / clip.w, -1.0f, 1.0f);
        const float ny = gclamp(clip.y / clip.w, -1.0f, 1.0f);
        const float u = nx * 0.5f + 0.5f, v = ny * 0.5f + 0.5f, z = clip.z / clip.w;
        mnx = gmin(mnx, u); mny = gmin(mny, v); mnz = gmin(mnz, z);
        mxx = gmax(mxx, u); mxy = gmax(mxy, v); mxz = gmax(mxz, z);
    }
    (void)mnz;
    const float width = (mxx - mnx) * w0;
    const float height = (mxy - mny) * h0;
    const float level = floor_log2(gmax(width, height));
    const float u = (mnx + mxx) * 0.5f, v = (mny + mxy) * 0.5f;
    float lam = level;
    if (!(lam >= 0.0f)) lam = 0.0f;
    if (lam > 16.0f) lam = 16.0f;
    uint32_t d = (uint32_t)(ceilf(lam + 0.5f) - 1.0f);
    if (d > last) d = last;
    ProbeAt p;
    p.lw = __shfl(pw, (int)d);
    const int lh = __shfl(ph, (int)d);
    p.off = (uint32_t)__shfl((int)po, (int)d);
    axis_footprint(u * (float)p.lw - 0.5f, p.lw, p.x0, p.x1);
    axis_footprint(v * (float)lh - 0.5f, lh, p.y0, p.y1);
    p.maxz = mxz;
    return p;
}

// The probe's depth: the minimum of its 2x2 footprint.
__device__ __forceinline__ float probe_fetch(const float* __restrict__ pyr, const ProbeAt& p)
{
    const float* tex = pyr + p.off;
    float r = tex[p.y0 * p.lw + p.x0];
    r = gmin(r, tex[p.y0 * p.lw + p.x1]);
    r = gmin(r, tex[p.y1 * p.lw + p.x0]);
    r = gmin(r, tex[p.y1 * p.lw + p.x1]);
    return r;
}

__device__ __forceinline__ float3 mat4_dir(const Mat4& m, float x, float y, float z)
{
    float3 o;
    o.x = ((m.m[0] * x + m.m[4] * y) + m.m[8] * z) + m.m[12] * 0.0f;
    o.y = ((m.m[1] * x + m.m[5] * y) + m.m[9] * z) + m.m[13] * 0.0f;
    o.z = ((m.m[2] * x + m.m[6] * y) + m.m[10] * z) + m.m[14] * 0.0f;
    return o;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}

// Normal-cone back-face test (meshoptimizer's published cone test), skipped
// unless the instance transform is a proper similarity.  The part that does
// not depend on the eye: the cone in world space (apex, unit axis), or false
// when the test does not apply.
__device__ __forceinline__ bool cone_setup(const vkgv_meshlet_cone& cone, const Mat4& m, float apex_out[3],
                                           float axis_out[3], float& cutoff)
{
    if (!(cone.cutoff < 1.0f)) return false;
    const float* a = m.m;
    const float l0 = dot3(a[0], a[1], a[2], a[0], a[1], a[2]);
    const float l1 = dot3(a[4], a[5], a[6], a[4], a[5], a[6]);
    const float l2 = dot3(a[8], a[9], a[10], a[8], a[9], a[10]);
    const float tol = l0 * 1e-5f;
    if (!(fabsf(l0 - l1) <= tol && fabsf(l0 - l2) <= tol && fabsf(dot3(a[0], a[1], a[2], a[4], a[5], a[6])) <= tol
          && fabsf(dot3(a[0], a[1], a[2], a[8], a[9], a[10])) <= tol
          && fabsf(dot3(a[4], a[5], a[6], a[8], a[9], a[10])) <= tol))
        return false;
    if (!(det4(m) > 0.0f)) return false;
    const float4 apex = mat4_point(m, cone.apex[0], cone.apex[1], cone.apex[2]);
    float3 ax = mat4_dir(m, cone.axis[0], cone.axis[1], cone.axis[2]);
    const float al = sqrtf(dot3(ax.x, ax.y, ax.z, ax.x, ax.y, ax.z));
    ax.x = ax.x / al; ax.y = ax.y / al; ax.z = ax.z / al;
    apex_out[0] = apex.x; apex_out[1] = apex.y; apex_out[2] = apex.z;
    axis_out[0] = ax.x; axis_out[1] = ax.y; axis_out[2] = ax.z;
    cutoff = cone.cutoff;
    return true;
}

// The eye-dependent part: back-facing for this eye.
__device__ __forceinline__ bool cone_culled_from(const float apex[3], const float axis[3], float cutoff,
                                                 const float* eye)
{
    float dx = apex[0] - eye[0], dy = apex[1] - eye[1], dz = apex[2] - eye[2];
    const float dl = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
    dx = dx / dl; dy = dy / dl; dz = dz / dl;
    return dot3(dx, dy, dz, axis[0], axis[1], axis[2]) >= cutoff;
}

struct CullArgs {
    SceneDev s;
    BatchDev b;
    uint32_t flags;
    uint32_t mode;
    uint32_t list;
    uint32_t units;          // work units: chunks * groups
    uint32_t groups;         // view groups per chunk
    uint32_t group_views;    // views per group (the last group may hold fewer)
};

// The view-independent part of a draw's cull: its world-space AABB and, for
// cone culling, the transformed cone.  Computed once per draw and reused for
// every view of the unit's view group.
struct DrawCull {
    float wc[3], we[3];
    float apex[3], axis[3], cutoff;
    bool valid, cone;
};

template <bool CONE>
__device__ __forceinline__ void draw_cull_setup(const SceneDev& s, uint32_t d, DrawCull& dc)
{
    dc.valid = d < s.draw_count;
    dc.cone = false;
    if (!dc.valid) {   // finite inputs for the probe every lane computes
        dc.wc[0] = dc.wc[1] = dc.wc[2] = 0.0f;
        dc.we[0] = dc.we[1] = dc.we[2] = 0.0f;
        return;
    }
    const float4* rp = reinterpret_cast<const float4*>(s.cull + d);
    const float4 r0 = rp[0], r1 = rp[1];
    const uint32_t tidx = __float_as_uint(r1.z);
    const Mat4 m = load_mat4(s.transforms + 16ull * tidx);
    // world AABB: centre (M * vec4(c,1)).xyz, extent mat3(|M0|,|M1|,|M2|) * e
    const float4 wc4 = mat4_point(m, r0.x, r0.y, r0.z);
    dc.wc[0] = wc4.x; dc.wc[1] = wc4.y; dc.wc[2] = wc4.z;
    const float e[3] = { r0.w, r1.x, r1.y };
#pragma unroll
    for (int r = 0; r < 3; ++r)
        dc.we[r] = (fabsf(m.m[r]) * e[0] + fabsf(m.m[4 + r]) * e[1]) + fabsf(m.m[8 + r]) * e[2];
    if (CONE) {
        const uint32_t ci = __float_as_uint(r1.w);
        if (ci != 0xFFFFFFFFu) dc.cone = cone_setup(s.cones[ci], m, dc.apex, dc.axis, dc.cutoff);
    }
}

// Persistent cull + compaction.  A work unit is 64 * P consecutive draws
// (lane l takes draws base + 64 i + l, i < P, so each of the P record loads
// of a wave is one contiguous 2 KiB span) and a group of at most 64 / P
// views.  Waves stride over the units.
//  - Per view, the P draws' tests run together: the occlusion probes' texel
//    loads are issued before any is compared, and the pyramid's level sizes
//    and offsets come from lanes (lane k holds level k & 15) by a cross-lane
//    read instead of a per-lane table load.
//  - Phase 1 records each lane's visibility per (view, draw) in a bitset and
//    reserves the unit's span of each view's list with one atomicAdd, issued
//    by lane j for view j and not waited on until phase 2, which writes the
//    raster records (draw order within the unit).
//  - The late pass writes the history words from the same ballots.
template <int P, bool CONE>
__global__ __launch_bounds__(kBlock) void cull_kernel(CullArgs a)
{
    static_assert(64 % P == 0, "a unit's (view, draw) bits fill at most one 64-bit word");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nwaves = (gridDim.x * kBlock) >> 6;
    const bool single = a.mode == CULL_SINGLE;
    const bool probe = a.mode == CULL_LATE || (single && (a.flags & VKGV_CULL_OCCLUSION_PREV));
    // the wave's index, uniform in the compiler's eyes: unit, view group and
    // view loop bounds then live in scalar registers and the loops branch on them
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * kBlock + threadIdx.x) >> 6);
    const uint32_t w = threadIdx.x >> 6;   // wave of the block
    __shared__ float s_dc[kBlock / 64][6][64 * P];              // world AABB centre, extent of the unit's draws
    __shared__ uint16_t s_pairs[kBlock / 64][64 * P];           // probe pairs: i << 6 | lane
    __shared__ unsigned long long s_occ[kBlock / 64][P];        // probe results per slot i
    __shared__ unsigned long long s_bits[kBlock / 64][kReserveChunks][64];   // a super-unit's chunks' bits
    __shared__ uint32_t s_off[kBlock / 64][kReserveChunks][64];              // and per-view offsets (lane j)
    if (a.mode != CULL_LATE) {
        // this frame's VP M and det(M) of every transform, per view, read by
        // the raster and the resolve (the same arithmetic they used per
        // meshlet and per pixel); the grid's threads share the (view,
        // transform) entries, one each, so that a batch's views run side by
        // side rather than one after another on the same threads
        const uint32_t gt = blockIdx.x * kBlock + threadIdx.x, nt = gridDim.x * kBlock;
        const uint32_t tc = a.s.transform_count;
        const unsigned long long entries = (unsigned long long)tc * a.b.nviews;
        for (unsigned long long e = gt; e < entries; e += nt) {
            const uint32_t vi = (uint32_t)(e / tc), t = (uint32_t)(e - (unsigned long long)vi * tc);
            const ViewDev& v = batch_view(a.b, vi);
            if (t >= v.mvp_capacity) continue;
            const Mat4 vp = load_mat4(batch_camera(a.b, vi).cam.view_projection);
            const Mat4 m = load_mat4(a.s.transforms + 16ull * t);
            const Mat4 mvp = mat4_mul(vp, m);
            float4* o = reinterpret_cast<float4*>(v.mvp + (size_t)t * kMvpStride);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                o[c] = make_float4(mvp.m[4 * c], mvp.m[4 * c + 1], mvp.m[4 * c + 2], mvp.m[4 * c + 3]);
            o[4] = make_float4(det4(m), 0.0f, 0.0f, 0.0f);
        }
    }
    if (a.mode == CULL_LATE && wave == 0) {
        // the first raster pass's queue counters, zeroed for the late pass
        // (their last reader, that pass's tile kernel, has finished)
        for (uint32_t v = lane; v < a.b.nviews; v += 64u) {
            Counters* c = batch_view(a.b, v).counters;
            c->big_chunks = 0ull;
            c->big_count = 0u;
            c->redo_count = 0u;
            c->clip_count = 0u;
            c->frag_jobs = 0ull;
        }
    }
    // A wave takes K consecutive chunks of one view group at a time (a
    // super-unit) and reserves their span of each view's list with one
    // atomicAdd: the views' counters see 1/K of the same-address atomics
    // (which the memory side serializes) they would with one per chunk.  K
    // is 1 while the units do not outnumber the waves (every wave keeps
    // work), up to kReserveChunks when they do.  Phase 1 keeps each chunk's
    // bits and per-view offsets in LDS.
    const uint32_t chunks = a.units / a.groups;
    const uint32_t K = max(1u, min(kReserveChunks, a.units / nwaves));
    const uint32_t supers = (chunks + K - 1u) / K * a.groups;
    for (uint32_t su = wave; su < supers; su += nwaves) {
        const uint32_t sc = su / a.groups, group = su - sc * a.groups;
        const uint32_t v0 = group * a.group_views, nv = min(a.group_views, a.b.nviews - v0);
        const uint32_t kn = min(K, chunks - sc * K);
        uint32_t run = 0;   // lane j: positions of the super-unit's chunks so far in view v0 + j
        for (uint32_t kc = 0; kc < kn; ++kc) {
            const uint32_t chunk = sc * K + kc;
            const uint32_t base = chunk * 64u * (uint32_t)P;
            DrawCull dc[P];
#pragma unroll
            for (int i = 0; i < P; ++i) draw_cull_setup<CONE>(a.s, base + 64u * i + lane, dc[i]);
            if (probe) {   // the world AABBs for the packed probes (the previous unit's reads are done)
#pragma unroll
                for (int i = 0; i < P; ++i)
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        s_dc[w][q][64 * i + lane] = dc[i].wc[q];
                        s_dc[w][3 + q][64 * i + lane] = dc[i].we[q];
                    }
            }
            unsigned long long bits = 0ull;   // bit j * P + i: draw i of this lane visible in view v0 + j
            uint32_t off = 0;                 // lane j: the chunk's first position in view v0 + j, from the reservation
            // the unit's history words of every view of the group in one load: lane
            // j * P + i holds view v0 + j's word of draws base + 64 i .. + 63 (read
            // before the late pass below rewrites them)
            unsigned long long hword = 0ull;
            if (!single) {
                const uint32_t jl = lane / (uint32_t)P, il = lane % (uint32_t)P;
                if (jl < nv && base + 64u * il < a.s.draw_count)
                    hword = reinterpret_cast<const unsigned long long*>(batch_view(a.b, v0 + jl).history)
                        [(base >> 6) + il];
            }
            // ---- phase 1: visibility of every (view, draw), one reservation per view
            for (uint32_t j = 0; j < nv; ++j) {
                const ViewDev& view = batch_view(a.b, v0 + j);
                const CameraArg& cam = batch_camera(a.b, v0 + j);
                unsigned long long* history = reinterpret_cast<unsigned long long*>(view.history);
                bool pass[P], vis[P];
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    // (readlane returns a signed int: widen each half as uint32_t, or a
                    // set bit 31 of the low half would sign-extend into the high half)
                    const uint32_t src = j * (uint32_t)P + (uint32_t)i;
                    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(hword >> 32), src);
                    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)hword, src);
                    const unsigned long long hw = ((unsigned long long)hi << 32) | (unsigned long long)lo;
                    const bool old = dc[i].valid && ((hw >> lane) & 1ull);
                    // the early pass tests only what the history lists
                    pass[i] = dc[i].valid && (a.mode != CULL_EARLY || old);
                    if (pass[i] && (a.flags & VKGV_CULL_FRUSTUM))
                        pass[i] = aabb_in_frustum(dc[i].wc, dc[i].we, cam.cam);
                    if (CONE && pass[i] && dc[i].cone)
                        pass[i] = !cone_culled_from(dc[i].apex, dc[i].axis, dc[i].cutoff, cam.eye);
                    vis[i] = single ? pass[i] : (a.mode == CULL_EARLY ? pass[i] : old);   // late: old = drawn early
                }
                if (probe) {
                    // The probes of the (i, lane) pairs past the frustum test, packed
                    // into rounds of 64, one pair per lane: in most views most of a
                    // unit's draws fail the test, so the projected-AABB arithmetic
                    // runs once per pair that needs it instead of once per lane of
                    // every slot i with a pair.  The draws' world AABBs come from
                    // LDS, the results go back through a bitmask per slot.
                    uint32_t npass = 0;
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        const unsigned long long m = __ballot(pass[i]);
                        const uint32_t pos =
                            npass + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        if (pass[i]) s_pairs[w][pos] = (uint16_t)(((uint32_t)i << 6) | lane);
                        npass += (uint32_t)__popcll(m);
                    }
                    if (lane < (uint32_t)P) s_occ[w][lane] = 0ull;
                    wave_sync();
                    if (npass) {   // wave-uniform
                        // pyramid level geometry in lanes: lane k holds level k & 15
                        const uint32_t lk = lane & 15u;
                        const int pw = (int)view.pyr.w[lk], ph = (int)view.pyr.h[lk];
                        const uint32_t po = view.pyr.off[lk];
                        const uint32_t last = view.pyr.levels - 1u;
                        const float w0 = (float)view.pyr.w[0], h0 = (float)view.pyr.h[0];
                        const float* vpm =
                            single ? cam.cam.prev_occlusion_view_projection : cam.cam.occlusion_view_projection;
                        Mat4 vp;
#pragma unroll
                        for (int q = 0; q < 16; ++q) vp.m[q] = vpm[q];
                        for (uint32_t r = 0; r < npass; r += 64u) {
                            const bool live = r + lane < npass;
                            const uint32_t e = live ? (uint32_t)s_pairs[w][r + lane] : 0u;
                            const uint32_t k = (e >> 6) * 64u + (e & 63u);   // the draw's slot in the unit
                            // finite inputs for the lanes without a pair (every lane computes)
                            float c[3] = { 0.0f, 0.0f, 0.0f }, ex[3] = { 0.0f, 0.0f, 0.0f };
                            if (live) {
#pragma unroll
                                for (int q = 0; q < 3; ++q) {
                                    c[q] = s_dc[w][q][k];
                                    ex[q] = s_dc[w][3 + q][k];
                                }
                            }
                            const ProbeAt pr = probe_setup(c, ex, vp, w0, h0, pw, ph, po, last);
                            const float depth = probe_fetch(view.pyramid, pr);
                            if (live && depth < pr.maxz) atomicOr(&s_occ[w][e >> 6], 1ull << (e & 63u));
                        }
                        wave_sync();
                    }
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        const bool occ = pass[i] && ((s_occ[w][i] >> lane) & 1ull);
                        if (single) {
                            vis[i] = occ;
                        } else {
                            const bool old = vis[i];
                            vis[i] = occ && !(pass[i] && old);
                            // the history of the next frame: visible against this frame's pyramid
                            const unsigned long long hb = __ballot(occ);
                            if (lane == 0 && base + 64u * i < a.s.draw_count) history[(base >> 6) + i] = hb;
                        }
                    }
                    wave_sync();   // s_pairs and s_occ are rewritten for the next view
                }
                uint32_t total = 0;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    total += (uint32_t)__popcll(__ballot(vis[i]));
                    if (vis[i]) bits |= 1ull << (j * (uint32_t)P + (uint32_t)i);
                }
                if (lane == j) {
                    off = run;
                    run += total;
                }
            }
            s_bits[w][kc][lane] = bits;
            s_off[w][kc][lane] = off;
        }
        // lane j: the super-unit's first list position in view v0 + j
        uint32_t reserved = 0;
        if (lane < nv && run) reserved = atomicAdd(&batch_view(a.b, v0 + lane).counters->vis_count[a.list], run);
        wave_sync();
        for (uint32_t kc = 0; kc < kn; ++kc) {
            const unsigned long long bits = s_bits[w][kc][lane];
            if (__ballot(bits != 0ull) == 0ull) continue;
            const uint32_t base = (sc * K + kc) * 64u * (uint32_t)P;
            const uint32_t first = reserved + s_off[w][kc][lane];   // lane j: the chunk's position in view v0 + j
            // ---- phase 2: the raster records of the visible draws.  A draw's
            // record is the same for every view: each lane loads the records of its
            // draws visible in any view once, all P loads in flight together, then
            // stores them into every view's list that holds them.
            uint4 rec[P][2];
#pragma unroll
            for (int i = 0; i < P; ++i) {
                unsigned long long in_any = 0ull;
                for (uint32_t j = 0; j < nv; ++j) in_any |= (bits >> (j * (uint32_t)P + (uint32_t)i)) & 1ull;
                rec[i][0] = rec[i][1] = make_uint4(0u, 0u, 0u, 0u);
                if (in_any) {
                    const uint4* src = reinterpret_cast<const uint4*>(a.s.rrec + base + 64u * i + lane);
                    rec[i][0] = src[0];
                    rec[i][1] = src[1];
                }
            }
            for (uint32_t j = 0; j < nv; ++j) {
                uint32_t out = __builtin_amdgcn_readlane(first, j);
                RasterRec* dst = batch_view(a.b, v0 + j).visible[a.list];
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const bool mine = (bits >> (j * (uint32_t)P + (uint32_t)i)) & 1ull;
                    const unsigned long long m = __ballot(mine);
                    if (mine) {
                        const uint32_t rank =
                            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        uint4* o = reinterpret_cast<uint4*>(dst + out + rank);
                        o[0] = rec[i][0];
                        o[1] = rec[i][1];
                    }
                    out += (uint32_t)__popcll(m);
                }
            }
        }
        wave_sync();   // s_bits and s_off are rewritten by the next super-unit
    }
}

// Next-frame clear, for frames without a resolve (which does it otherwise):
// each block clears kClearKeysPerBlock keys of its view's next key buffer
// (kClearKeysPerBlock / kBlock = 32 per thread, two per 16-byte store); the
// view's first block also zeroes the next frame's counters, and the view's
// blocks zero the per-tile triangle counts of both raster passes.
static_assert(kClearKeysPerBlock % (2 * kBlock) == 0, "the unrolled clear stores whole key pairs per thread");
__global__ __launch_bounds__(kBlock) void next_frame_clear_kernel(BatchDev b)
{
    const uint32_t e = b.map_clear[blockIdx.x];
    const uint32_t vi = e & 0xFFu, local = e >> 8;
    const ViewDev& v = batch_view(b, vi);
    if (local == 0 && threadIdx.x < sizeof(Counters) / 4) reinterpret_cast<uint32_t*>(v.counters_next)[threadIdx.x] = 0u;
    const uint64_t n = (uint64_t)v.width * v.height;
    {
        const uint32_t nblk = (uint32_t)((n + kClearKeysPerBlock - 1) / kClearKeysPerBlock);
        const uint32_t ncount = 2u * tile_count_stride(v);
        for (uint32_t i = local * kBlock + threadIdx.x; i < ncount; i += nblk * kBlock) v.bin_count[i] = 0u;
    }
    const uint64_t first = (uint64_t)local * kClearKeysPerBlock;
    const ulonglong2 c2 = make_ulonglong2(VKGV_KEY_CLEAR, VKGV_KEY_CLEAR);
    ulonglong2* k2 = reinterpret_cast<ulonglong2*>(v.keys_next + first);
    if (first + kClearKeysPerBlock <= n) {
#pragma unroll
        for (uint32_t q = 0; q < kClearKeysPerBlock / 2u / kBlock; ++q) k2[q * kBlock + threadIdx.x] = c2;
    } else {
        for (uint64_t i = first + threadIdx.x; i < n; i += kBlock) v.keys_next[i] = VKGV_KEY_CLEAR;
    }
}

// prevViewProjection * previous M of every transform, per view (the mesh
// shader's prevMvp, visbuffer.mesh.glsl:45, with the previous frame's
// transform where the reference reuses the current one).
__global__ __launch_bounds__(kBlock) void prev_mvp_kernel(SceneDev s, BatchDev b)
{
    const uint32_t vi = blockIdx.y;
    const ViewDev& v = batch_view(b, vi);
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= s.transform_count || t >= v.pmvp_capacity) return;
    const Mat4 pvp = load_mat4(batch_camera(b, vi).cam.prev_view_projection);
    const Mat4 m = mat4_mul(pvp, load_mat4(s.prev_transforms + 16ull * t));
    float4* o = reinterpret_cast<float4*>(v.pmvp + 16ull * t);
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = make_float4(m.m[4 * c], m.m[4 * c + 1], m.m[4 * c + 2], m.m[4 * c + 3]);
}

// Cull and raster records from the contract arrays: draw -> primitive ->
// meshlet AABB, buffer offsets, counts and material.
__global__ __launch_bounds__(kBlock) void build_cull_records_kernel(SceneDev s, CullRec* out, RasterRec* rrec)
{
    const uint32_t d = blockIdx.x * kBlock + threadIdx.x;
    if (d >= s.draw_count) return;
    const vkgv_meshlet_draw dr = s.draws[d];
    const vkgv_primitive& p = s.prims[dr.primitive_index];
    const vkgv_meshlet* ml = reinterpret_cast<const vkgv_meshlet*>(p.meshlet_buffer) + dr.meshlet_index;
    CullRec r;
    r.cx = ml->aabb_center[0]; r.cy = ml->aabb_center[1]; r.cz = ml->aabb_center[2];
    r.ex = ml->aabb_extents[0]; r.ey = ml->aabb_extents[1]; r.ez = ml->aabb_extents[2];
    r.transform = dr.transform_index;
    const uint32_t cb = s.prim_cone_base[dr.primitive_index];
    r.cone = cb == 0xFFFFFFFFu ? 0xFFFFFFFFu : cb + dr.meshlet_index;
    out[d] = r;
    RasterRec q;
    q.draw = d;
    q.transform = dr.transform_index;
    q.vidx = (uint32_t)((reinterpret_cast<const uint32_t*>(p.vertex_index_buffer) - s.vidx_base) + ml->vertex_offset);
    q.tri = (uint32_t)((reinterpret_cast<const uint8_t*>(p.primitive_index_buffer) - s.tri_base) + ml->triangle_offset);
    // byte offset: meshes are packed at 256-byte boundaries, not at multiples of a vertex
    q.vtx = (uint32_t)(reinterpret_cast<const uint8_t*>(p.vertex_buffer) - reinterpret_cast<const uint8_t*>(s.vtx_base));
    q.material = p.material_index | (s.materials[p.material_index].double_sided ? 0x80000000u : 0u);
    q.counts = (uint32_t)ml->vertex_count | ((uint32_t)ml->triangle_count << 8);
    q.pad = 0u;
    rrec[d] = q;
}

// One thread per meshlet of a scene uploaded from device memory: the same
// conditions the host upload checks (counts within the limits, 4-byte
// aligned triangle offset, ranges within the mesh's arrays, vertex indices
// below the vertex count, triangle indices below the meshlet's vertex count),
// read from the packed copies only after the ranges are known to hold.
__global__ __launch_bounds__(kBlock) void check_meshes_kernel(const MeshCheck* meshes, uint32_t mesh_count,
                                                              uint64_t meshlets, const uint32_t* vidx,
                                                              const uint8_t* tri, const uint8_t* ml,
                                                              unsigned long long* bad)
{
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= meshlets) return;
    uint32_t lo = 0, hi = mesh_count - 1;   // the last mesh whose first meshlet is <= g
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (meshes[mid].ml_first <= g) lo = mid; else hi = mid - 1u;
    }
    const MeshCheck m = meshes[lo];
    const uint64_t k = g - m.ml_first;
    const vkgv_meshlet* p = reinterpret_cast<const vkgv_meshlet*>(ml + m.ml_off) + k;
    const uint32_t voff = p->vertex_offset, toff = p->triangle_offset;
    const uint32_t vc = p->vertex_count, tc = p->triangle_count;
    const uint64_t tbytes = ((uint64_t)tc * 3u + 3u) & ~3ull;
    bool ok = vc >= 1u && vc <= VKGV_MAX_VERTICES && tc >= 1u && tc <= VKGV_MAX_PRIMITIVES && (toff & 3u) == 0u
              && (uint64_t)voff + vc <= m.vidx_count && (uint64_t)toff + tbytes <= m.tri_bytes;
    for (uint32_t v = 0; ok && v < vc; ++v) ok = vidx[m.vidx_off + voff + v] < m.vtx_count;
    for (uint32_t t = 0; ok && t < 3u * tc; ++t) ok = tri[m.tri_off + toff + t] < vc;
    if (!ok) atomicMin(bad, ((unsigned long long)lo << 32) | k);
}

__global__ __launch_bounds__(kBlock) void clear_keys_kernel(ulonglong2* keys, uint64_t n2, unsigned long long* tail, uint32_t odd)
{
    const ulonglong2 v = make_ulonglong2(VKGV_KEY_CLEAR, VKGV_KEY_CLEAR);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * kBlock) keys[i] = v;
    if (odd && blockIdx.x == 0 && threadIdx.x == 0) *tail = VKGV_KEY_CLEAR;
}

__global__ __launch_bounds__(kBlock) void write_depth_keys_kernel(unsigned long long* keys, const float* depth, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) keys[i] = ((unsigned long long)(~__float_as_uint(depth[i])) << 32) | 0xFFFFFFFFull;
}

} // namespace

void launch_build_cull_records(const SceneDev& s, CullRec* out, RasterRec* rrec, hipStream_t st)
{
    if (!s.draw_count) return;
    build_cull_records_kernel<<<(s.draw_count + kBlock - 1) / kBlock, kBlock, 0, st>>>(s, out, rrec);
}

hipError_t launch_check_meshes(const MeshCheck* meshes, uint32_t mesh_count, uint64_t meshlets, const uint32_t* vidx,
                               const uint8_t* tri, const uint8_t* ml, unsigned long long* bad, hipStream_t st)
{
    if (!meshlets || !mesh_count) return hipSuccess;
    const uint64_t blocks = (meshlets + kBlock - 1) / kBlock;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    check_meshes_kernel<<<(uint32_t)blocks, kBlock, 0, st>>>(meshes, mesh_count, meshlets, vidx, tri, ml, bad);
    return hipGetLastError();
}

void launch_clear_keys(unsigned long long* keys, uint64_t n, hipStream_t st)
{
    const uint64_t n2 = n / 2;
    uint64_t blocks = (n2 + kBlock - 1) / kBlock;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) blocks = 1;
    clear_keys_kernel<<<(uint32_t)blocks, kBlock, 0, st>>>(reinterpret_cast<ulonglong2*>(keys), n2, keys + n - 1,
                                                           (uint32_t)(n & 1));
}

void launch_next_frame_clear(const BatchDev& b, const BatchGrid& g, hipStream_t st)
{
    next_frame_clear_kernel<<<g.clear_blocks, kBlock, 0, st>>>(b);
}

void launch_prev_mvp(const SceneDev& s, const BatchDev& b, const BatchGrid& g, hipStream_t st)
{
    if (!s.transform_count) return;
    prev_mvp_kernel<<<dim3((s.transform_count + kBlock - 1) / kBlock, g.nviews), kBlock, 0, st>>>(s, b);
}

namespace {

// Persistent grid: as many blocks as stay resident (the kernel's occupancy,
// queried once per instantiation), each wave striding over the units; fewer
// when there are fewer units.
template <int P, bool CONE>
void launch_cull_kernel(const CullArgs& a, uint32_t cus, hipStream_t st)
{
    static const uint32_t per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, cull_kernel<P, CONE>, kBlock, 0) != hipSuccess || n < 1)
            n = 1;
        return (uint32_t)n;
    }();
    const uint32_t blocks = std::min<uint32_t>((a.units + 3u) / 4u, cus * per_cu);
    cull_kernel<P, CONE><<<blocks, kBlock, 0, st>>>(a);
}

} // namespace

void launch_cull(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, CullMode mode,
                 uint32_t list, uint32_t cus, hipStream_t st)
{
    if (!s.draw_count || !g.nviews) return;
    CullArgs a;
    a.s = s;
    a.b = b;
    a.flags = flags;
    a.mode = mode;
    a.list = list;
    // Four draws per lane from 2M draws on (fewer atomics and loop trips per
    // draw), one below (more waves for small lists).  Views are grouped so
    // that the units still number at least 32 per CU: a group shares each
    // record and transform load, and a many-view batch of a small scene gets
    // its parallelism from the groups.
    const uint32_t per_lane = s.draw_count >= (1u << 21) ? 4u : 1u;
    const uint32_t chunks = (s.draw_count + 64u * per_lane - 1) / (64u * per_lane);
    const uint32_t target = cus * 32u;
    uint32_t groups = std::min<uint32_t>(g.nviews, std::max<uint32_t>(1u, (target + chunks - 1) / chunks));
    // a lane keeps one visibility bit per (view, draw) of its unit: at most 64
    a.group_views = std::min<uint32_t>((g.nviews + groups - 1) / groups, 64u / per_lane);
    a.groups = (g.nviews + a.group_views - 1) / a.group_views;
    a.units = chunks * a.groups;
    const bool cone = flags & VKGV_CULL_CONE;
    if (per_lane == 4) {
        if (cone) launch_cull_kernel<4, true>(a, cus, st);
        else launch_cull_kernel<4, false>(a, cus, st);
    } else {
        if (cone) launch_cull_kernel<1, true>(a, cus, st);
        else launch_cull_kernel<1, false>(a, cus, st);
    }
}

void launch_write_depth_keys(unsigned long long* keys, const float* depth, uint64_t n, hipStream_t st)
{
    if (!n) return;
    write_depth_keys_kernel<<<(uint32_t)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(keys, depth, n);
}

} // namespace vkgv
