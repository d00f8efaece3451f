// Material resolve (visbuffer_resolve.comp.glsl:17-41, srgb.h.glsl:26-32).
//
// Per pixel: key -> id -> draw -> primitive -> material; sRGB-encode the
// albedo factor (alpha passes through) to RGBA8; the clear key gives (0,0,0,0).
// With VKGV_RESOLVE_BARY perspective-correct barycentrics are recomputed at
// the pixel centre from the triangle's clip-space vertices (2D homogeneous
// form, valid for clipped triangles too) to interpolate the vertex colour,
// which multiplies the albedo factor.
//
// The raster leaves each rasterized meshlet's albedo and, per triangle, the
// cross products of its clip-space (x, y, w) vertices and its vertex colours
// in a per-view cache indexed by the meshlet's slot in the visible list, so a
// pixel needs key -> slot -> triangle (three dependent loads); meshlets beyond
// the cache capacity are re-fetched and re-transformed from the scene buffers
// (the same arithmetic).  sRGB encode + unorm8 is a two-level table lookup
// (exact for the 8-bit result).
//
// VKGV_RESOLVE_TEXTURED adds the albedo texture the reference's fragment
// shader leaves commented out (visbuffer.frag.glsl:31-32): the vertex uvs
// (float16; kept in the resolve cache beside the triangle, with the slot's
// material) are interpolated with the same barycentrics, transformed by
// transformUv (mesh_common.h.glsl:129-135: rotation, then scale, then
// offset), and the texture is sampled bilinearly at level 0 with per-texel
// sRGB decode; colour = vertex colour * albedo factor * texture.
//
// VKGV_MOTION_VECTORS writes each pixel's motion vector beside the colour.
#include "internal.hpp"

namespace vkgv {
namespace {

__device__ __forceinline__ uint32_t unorm8(float x)
{
    float v = gclamp(x, 0.0f, 1.0f);
    if (!(v >= 0.0f)) v = 0.0f;
    return (uint32_t)(v * 255.0f + 0.5f);
}

// unorm8(fromLinear(c)): 0 below the first threshold (and for NaN), 255 from
// 1.0 up, else base[bucket] plus the one threshold the bucket may hold.
__device__ __forceinline__ uint32_t srgb_code(const float* thr, const uint8_t* base, uint32_t key0, float c)
{
    if (!(c >= thr[0])) return 0u;
    if (c >= 1.0f) return 255u;
    const uint32_t k = base[(__float_as_uint(c) >> 16) - key0];
    return k + (c >= thr[k] ? 1u : 0u);
}

struct ResolveArgs {
    SceneDev s;
    const unsigned long long* keys;
    uint32_t* color;
    uint32_t W, H;
    uint32_t flags;
    const uint32_t* draw_slot;
    const TriCache* tri_cache;
    const float4* slot_albedo;
    const uint32_t* slot_mat;
    const float* mvp;               // the view's VP M table (kMvpStride floats per transform)
};

__device__ __forceinline__ float3 cross3(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Albedo and triangle data of a draw whose meshlet is not in the resolve
// cache: re-fetched from the scene and re-transformed (same arithmetic as the
// raster's cache fill).
__device__ __forceinline__ void fetch_uncached(const ResolveArgs& a, uint32_t d, uint32_t t, bool bary, float* c, float3* r,
                                               uint32_t* col)
{
    const vkgv_meshlet_draw dr = a.s.draws[d];
    const vkgv_primitive& prim = a.s.prims[dr.primitive_index];
    const vkgv_material& mat = a.s.materials[prim.material_index];
    c[0] = mat.albedo_factor[0]; c[1] = mat.albedo_factor[1]; c[2] = mat.albedo_factor[2]; c[3] = mat.albedo_factor[3];
    if (!bary) return;
    const vkgv_meshlet ml = reinterpret_cast<const vkgv_meshlet*>(prim.meshlet_buffer)[dr.meshlet_index];
    const Mat4 mvp = load_mat4(a.mvp + (size_t)dr.transform_index * kMvpStride);   // VP M of this frame
    const uint8_t* tri = reinterpret_cast<const uint8_t*>(prim.primitive_index_buffer) + ml.triangle_offset + 3u * t;
    const uint32_t* vidx = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer) + ml.vertex_offset;
    const vkgv_vertex* verts = reinterpret_cast<const vkgv_vertex*>(prim.vertex_buffer);
    float3 hv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const vkgv_vertex& vx = verts[vidx[tri[k]]];
        const float4 cl = mat4_point(mvp, vx.position[0], vx.position[1], vx.position[2]);
        hv[k] = make_float3(cl.x, cl.y, cl.w);
        col[k] = *reinterpret_cast<const uint32_t*>(vx.color);
    }
    r[0] = cross3(hv[1], hv[2]);
    r[1] = cross3(hv[2], hv[0]);
    r[2] = cross3(hv[0], hv[1]);
}

// float16 bits -> float (exact: every half is a float).
__device__ __forceinline__ float half_bits_to_float(uint32_t h)
{
    const uint32_t sgn = (h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t bits;
    if (e == 0u) {
        const float f = (float)m * 5.9604644775390625e-8f;     // m * 2^-24
        return sgn ? -f : f;
    }
    if (e == 31u) bits = sgn | 0x7F800000u | (m << 13);
    else bits = sgn | ((e + 112u) << 23) | (m << 13);
    return __uint_as_float(bits);
}

// Texel index of integer coordinate k under a wrap mode.
__device__ __forceinline__ uint32_t wrap_texel(long long k, uint32_t n, uint32_t mode)
{
    if (mode == VKGV_WRAP_CLAMP_TO_EDGE) return (uint32_t)(k < 0 ? 0 : (k >= (long long)n ? n - 1 : k));
    if (mode == VKGV_WRAP_MIRRORED_REPEAT) {
        long long m = k % (2ll * n);
        if (m < 0) m += 2ll * n;
        return (uint32_t)(m < (long long)n ? m : 2ll * n - 1 - m);
    }
    long long r = k % (long long)n;
    return (uint32_t)(r < 0 ? r + n : r);
}

// Bilinear footprint along one axis: texels i0, i1 and the weight of i1.
// Coordinates are clamped to +-2^30 texels (beyond, the fraction is 0 anyway).
__device__ __forceinline__ float texel_pair(float u, uint32_t n, uint32_t mode, uint32_t& i0, uint32_t& i1)
{
    float x = u * (float)n - 0.5f;
    if (!(x == x)) x = 0.0f;
    x = x < -1073741824.0f ? -1073741824.0f : (x > 1073741824.0f ? 1073741824.0f : x);
    const float fl = floorf(x);
    const long long k = (long long)fl;
    i0 = wrap_texel(k, n, mode);
    i1 = wrap_texel(k + 1, n, mode);
    return x - fl;
}

// Material and float16 uv words of triangle t of draw d from the scene
// buffers (meshlets the resolve cache does not hold).
__device__ __forceinline__ uint32_t fetch_uvs(const SceneDev& s, uint32_t d, uint32_t t, uint32_t* uvw)
{
    const vkgv_meshlet_draw dr = s.draws[d];
    const vkgv_primitive& prim = s.prims[dr.primitive_index];
    const vkgv_meshlet ml = reinterpret_cast<const vkgv_meshlet*>(prim.meshlet_buffer)[dr.meshlet_index];
    const uint8_t* tri = reinterpret_cast<const uint8_t*>(prim.primitive_index_buffer) + ml.triangle_offset + 3u * t;
    const uint32_t* vidx = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer) + ml.vertex_offset;
    const vkgv_vertex* verts = reinterpret_cast<const vkgv_vertex*>(prim.vertex_buffer);
#pragma unroll
    for (int k = 0; k < 3; ++k) uvw[k] = *reinterpret_cast<const uint32_t*>(verts[vidx[tri[k]]].uv);
    return prim.material_index;
}

// Albedo texture of material mi times the pixel's colour (c[0..3]): the
// vertex uvs interpolated with l[], transformUv, bilinear sample, multiply.
__device__ __forceinline__ void apply_texture(const SceneDev& s, uint32_t mi, const uint32_t* uvw, const float* l,
                                              float* c)
{
    const MatTex mt = s.mat_tex[mi];
    if (mt.width == 0u) return;   // no texture
    float uq[3], vq[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uq[k] = half_bits_to_float(uvw[k] & 0xFFFFu);
        vq[k] = half_bits_to_float(uvw[k] >> 16);
    }
    const float u = (l[0] * uq[0] + l[1] * uq[1]) + l[2] * uq[2];
    const float v = (l[0] * vq[0] + l[1] * vq[1]) + l[2] * vq[2];
    // transformUv: mat2(cos, -sin, sin, cos) * uv * scale + offset
    const float ur = mt.cos_r * u + mt.sin_r * v;
    const float vr = (-mt.sin_r) * u + mt.cos_r * v;
    const float tu = ur * mt.uv_scale[0] + mt.uv_offset[0];
    const float tv = vr * mt.uv_scale[1] + mt.uv_offset[1];
    uint32_t x0, x1, y0, y1;
    const float fx = texel_pair(tu, mt.width, mt.wrap_s, x0, x1);
    const float fy = texel_pair(tv, mt.height, mt.wrap_t, y0, y1);
    const float4* img = s.texels + mt.offset;
    const float4 t00 = img[(size_t)y0 * mt.width + x0], t10 = img[(size_t)y0 * mt.width + x1];
    const float4 t01 = img[(size_t)y1 * mt.width + x0], t11 = img[(size_t)y1 * mt.width + x1];
    const float ax = 1.0f - fx, ay = 1.0f - fy;
    const float a00[4] = { t00.x, t00.y, t00.z, t00.w }, a10[4] = { t10.x, t10.y, t10.z, t10.w };
    const float a01[4] = { t01.x, t01.y, t01.z, t01.w }, a11[4] = { t11.x, t11.y, t11.z, t11.w };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float r0 = a00[j] * ax + a10[j] * fx;
        const float r1 = a01[j] * ax + a11[j] * fx;
        c[j] = c[j] * (r0 * ay + r1 * fy);
    }
}

// Motion vector of a pixel (visbuffer.frag.glsl:38): the triangle's vertices
// through prevViewProjection * previous M, interpolated with the pixel's
// perspective-correct barycentrics l[], against the pixel centre's NDC.
__device__ __forceinline__ float2 motion_vector(const SceneDev& s, const float* pmvp, uint32_t d, uint32_t t,
                                                const float* l, float xn, float yn)
{
    const vkgv_meshlet_draw dr = s.draws[d];
    const vkgv_primitive& prim = s.prims[dr.primitive_index];
    const vkgv_meshlet ml = reinterpret_cast<const vkgv_meshlet*>(prim.meshlet_buffer)[dr.meshlet_index];
    const uint8_t* tri = reinterpret_cast<const uint8_t*>(prim.primitive_index_buffer) + ml.triangle_offset + 3u * t;
    const uint32_t* vidx = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer) + ml.vertex_offset;
    const vkgv_vertex* verts = reinterpret_cast<const vkgv_vertex*>(prim.vertex_buffer);
    const Mat4 pm = load_mat4(pmvp + 16ull * dr.transform_index);
    float4 q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const vkgv_vertex& vx = verts[vidx[tri[k]]];
        q[k] = mat4_point(pm, vx.position[0], vx.position[1], vx.position[2]);
    }
    const float px = (l[0] * q[0].x + l[1] * q[1].x) + l[2] * q[2].x;
    const float py = (l[0] * q[0].y + l[1] * q[1].y) + l[2] * q[2].y;
    const float pw = (l[0] * q[0].w + l[1] * q[1].w) + l[2] * q[2].w;
    return make_float2(((px / pw) * 0.5f - 0.5f) - (xn * 0.5f - 0.5f),
                       ((py / pw) * 0.5f - 0.5f) - (yn * 0.5f - 0.5f));
}

// Block = 64 x (4 * kResolveRows) pixels; thread (x, y) resolves rows
// y, y + 4, ... of its column, with the loads of its pixels issued together.
constexpr int kRows = (int)kResolveRows;

struct ResolveBatchArgs {
    SceneDev s;
    BatchDev b;
    uint32_t flags;
};

// One block per 64x16 pixels of a view (block map entry = view | block << 8).
// One instantiation per shading mode, so that the flat and barycentric
// resolves carry none of the texture / motion-vector code: kShade =
// interpolated vertex colour, kTex = albedo texture, kMv = motion vectors.
template <bool kShade, bool kTex, bool kMv>
__global__ __launch_bounds__(256) void resolve_kernel(ResolveBatchArgs ab)
{
    const uint32_t me = ab.b.map_resolve[blockIdx.x];
    const ViewDev& view = batch_view(ab.b, me & 0xFFu);
    const uint32_t local = me >> 8;
    ResolveArgs a;
    a.s = ab.s;
    a.keys = view.keys;
    a.color = view.color;
    a.W = view.width;
    a.H = view.height;
    a.flags = ab.flags;
    a.draw_slot = view.draw_slot;
    a.tri_cache = view.tri_cache;
    a.slot_albedo = view.slot_albedo;
    a.slot_mat = view.slot_mat;
    a.mvp = view.mvp;
    float2* const view_motion = view.motion;
    const float* const view_pmvp = view.pmvp;
    const uint32_t blk_x = local % view.resolve_blocks_x, blk_y = local / view.resolve_blocks_x;
    // the next frame's state: its counters (first block) and the per-tile
    // triangle counts of both raster passes (spread over the view's blocks)
    {
        if (local == 0 && threadIdx.x < sizeof(Counters) / 4)
            reinterpret_cast<uint32_t*>(view.counters_next)[threadIdx.x] = 0u;
        const uint32_t nblk = view.resolve_blocks_x * ((view.height + 4u * kRows - 1u) / (4u * kRows));
        const uint32_t ncount = 2u * tile_count_stride(view);
        for (uint32_t i = local * 256u + threadIdx.x; i < ncount; i += nblk * 256u) view.bin_count[i] = 0u;
    }

    // LDS tables: sRGB thresholds and buckets, exact quotients b / 255 of a byte
    __shared__ float s_thr[256];
    __shared__ float s_byte[256];
    __shared__ uint32_t s_base[512];
    constexpr bool tex = kTex;
    const SrgbTables* tab = static_cast<const SrgbTables*>(a.s.srgb);
    s_thr[threadIdx.x] = tab->thr[threadIdx.x];
    s_byte[threadIdx.x] = (float)threadIdx.x / 255.0f;
    s_base[threadIdx.x] = reinterpret_cast<const uint32_t*>(tab->base)[threadIdx.x];
    s_base[threadIdx.x + 256] = reinterpret_cast<const uint32_t*>(tab->base)[threadIdx.x + 256];
    const uint32_t key0 = tab->key0;
    __syncthreads();
    const uint8_t* base = reinterpret_cast<const uint8_t*>(s_base);

    const uint32_t px = blk_x * 64u + (threadIdx.x & 63u);
    if (px >= a.W) return;
    const uint32_t py0 = blk_y * (4u * kRows) + (threadIdx.x >> 6);
    constexpr bool shade = kShade;                 // interpolated colour
    constexpr bool mv = kMv;
    constexpr bool bary = kShade || kMv;           // barycentrics needed
    const unsigned long long* __restrict__ keys = a.keys;
    const uint32_t* __restrict__ draw_slot = a.draw_slot;

    unsigned long long key[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint32_t py = py0 + 4u * k;
        key[k] = py < a.H ? keys[(size_t)py * a.W + px] : VKGV_KEY_CLEAR;
    }
    // the next frame's key buffer, cleared where this frame's keys are read
    // (the frame-start clear, folded into a pass that is not bound by HBM)
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint32_t py = py0 + 4u * k;
        if (py < a.H) view.keys_next[(size_t)py * a.W + px] = VKGV_KEY_CLEAR;
    }
    // draw of each pixel; ~0u: no shading (clear key or a key not written by this draw list)
    uint32_t drw[kRows], slot[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        drw[k] = 0xFFFFFFFFu;
        slot[k] = 0xFFFFFFFFu;
        if (key[k] != VKGV_KEY_CLEAR) {
            const uint32_t d = key_id(key[k]) >> VKGV_TRIANGLE_BITS;
            if (d < a.s.draw_count) { drw[k] = d; slot[k] = draw_slot[d]; }
        }
    }
    const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
    const float xn = (((float)px + 0.5f) - hw) * (1.0f / hw);
#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const uint32_t py = py0 + 4u * k;
        if (py >= a.H) break;
        const size_t i = (size_t)py * a.W + px;
        if (drw[k] == 0xFFFFFFFFu) {
            a.color[i] = 0u;
            if constexpr (mv) view_motion[i] = make_float2(0.0f, 0.0f);
            continue;
        }
        const uint32_t t = key_id(key[k]) & ((1u << VKGV_TRIANGLE_BITS) - 1u);
        float c[4];
        float3 r[3];
        uint32_t col[3], uvw[3], mi = 0;
        if (slot[k] != 0xFFFFFFFFu) {
            // the raster kept this meshlet's albedo and triangles
            const float4 alb = a.slot_albedo[slot[k]];
            c[0] = alb.x; c[1] = alb.y; c[2] = alb.z; c[3] = alb.w;
            if (bary) {
                const TriCache tc = a.tri_cache[(size_t)slot[k] * 128u + t];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    r[q] = make_float3(tc.r[q][0], tc.r[q][1], tc.r[q][2]);
                    col[q] = tc.col[q];
                    if constexpr (tex) uvw[q] = tc.uv[q];
                }
                if constexpr (tex) mi = a.slot_mat[slot[k]];
            }
        } else {
            fetch_uncached(a, drw[k], t, bary, c, r, col);
            if constexpr (tex) mi = fetch_uvs(a.s, drw[k], t, uvw);
        }
        if (bary) {
            // perspective-correct barycentrics at the pixel centre:
            // lambda_q ~ r[q] . (xn, yn, 1), normalized; c *= interpolated colour
            const float yn = (((float)py + 0.5f) - hh) * (1.0f / hh);
            float l[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) l[q] = (r[q].x * xn + r[q].y * yn) + r[q].z;
            const float inv = 1.0f / ((l[0] + l[1]) + l[2]);
#pragma unroll
            for (int q = 0; q < 3; ++q) l[q] = l[q] * inv;
            if (shade) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float c0 = s_byte[(col[0] >> (8 * j)) & 0xFFu];
                    const float c1 = s_byte[(col[1] >> (8 * j)) & 0xFFu];
                    const float c2 = s_byte[(col[2] >> (8 * j)) & 0xFFu];
                    c[j] = c[j] * ((l[0] * c0 + l[1] * c1) + l[2] * c2);
                }
            }
            if constexpr (tex) apply_texture(a.s, mi, uvw, l, c);
            if constexpr (mv) view_motion[i] = motion_vector(a.s, view_pmvp, drw[k], t, l, xn, yn);
        }
        const uint32_t r8 = srgb_code(s_thr, base, key0, c[0]);
        const uint32_t g8 = srgb_code(s_thr, base, key0, c[1]);
        const uint32_t b8 = srgb_code(s_thr, base, key0, c[2]);
        a.color[i] = r8 | (g8 << 8) | (b8 << 16) | (unorm8(c[3]) << 24);
    }
}

} // namespace

void launch_resolve(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, hipStream_t st)
{
    ResolveBatchArgs a;
    a.s = s;
    a.b = b;
    a.flags = flags;
    if (!g.resolve_blocks) return;
    const bool tex = flags & VKGV_RESOLVE_TEXTURED, mv = flags & VKGV_MOTION_VECTORS;
    const bool shade = tex || (flags & VKGV_RESOLVE_BARY);
    if (mv) {
        if (tex) resolve_kernel<true, true, true><<<g.resolve_blocks, 256, 0, st>>>(a);
        else if (shade) resolve_kernel<true, false, true><<<g.resolve_blocks, 256, 0, st>>>(a);
        else resolve_kernel<false, false, true><<<g.resolve_blocks, 256, 0, st>>>(a);
    } else {
        if (tex) resolve_kernel<true, true, false><<<g.resolve_blocks, 256, 0, st>>>(a);
        else if (shade) resolve_kernel<true, false, false><<<g.resolve_blocks, 256, 0, st>>>(a);
        else resolve_kernel<false, false, false><<<g.resolve_blocks, 256, 0, st>>>(a);
    }
}

} // namespace vkgv
