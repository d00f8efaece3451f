// Material resolve (visbuffer_resolve.comp.glsl:17-41, srgb.h.glsl:26-32).
//
// Per pixel: key -> id -> draw -> primitive -> material; sRGB-encode the
// albedo factor (alpha passes through) to RGBA8; the clear key gives (0,0,0,0).
// With VKGV_RESOLVE_BARY the triangle is re-fetched from the id, its vertices
// re-transformed, and perspective-correct barycentrics recomputed at the pixel
// centre (2D homogeneous form, valid for clipped triangles too) to interpolate
// the vertex colour, which multiplies the albedo factor.
#include "internal.hpp"

namespace vkgv {
namespace {

__device__ __forceinline__ float from_linear(float c)
{
    const bool cutoff = c < 0.0031308f;
    const float higher = 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    const float lower = c * 12.92f;
    return cutoff ? lower : higher;
}

__device__ __forceinline__ uint32_t unorm8(float x)
{
    float v = gclamp(x, 0.0f, 1.0f);
    if (!(v >= 0.0f)) v = 0.0f;
    return (uint32_t)(v * 255.0f + 0.5f);
}

struct ResolveArgs {
    SceneDev s;
    Mat4 vp;
    const unsigned long long* keys;
    uint32_t* color;
    uint32_t W, H;
    uint32_t flags;
};

__device__ __forceinline__ float3 cross3(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__global__ __launch_bounds__(256) void resolve_kernel(ResolveArgs a)
{
    const uint32_t px = blockIdx.x * 64u + (threadIdx.x & 63u);
    const uint32_t py = blockIdx.y * 4u + (threadIdx.x >> 6);
    if (px >= a.W || py >= a.H) return;
    const size_t i = (size_t)py * a.W + px;
    const unsigned long long key = a.keys[i];
    if (key == VKGV_KEY_CLEAR) { a.color[i] = 0u; return; }
    const uint32_t id = key_id(key);
    const uint32_t d = id >> VKGV_TRIANGLE_BITS, t = id & ((1u << VKGV_TRIANGLE_BITS) - 1u);
    if (d >= a.s.draw_count) { a.color[i] = 0u; return; }   // keys not written by this draw list
    const vkgv_meshlet_draw dr = a.s.draws[d];
    const vkgv_primitive& prim = a.s.prims[dr.primitive_index];
    const vkgv_material& mat = a.s.materials[prim.material_index];
    float c[4] = { mat.albedo_factor[0], mat.albedo_factor[1], mat.albedo_factor[2], mat.albedo_factor[3] };
    if (a.flags & VKGV_RESOLVE_BARY) {
        const vkgv_meshlet ml = reinterpret_cast<const vkgv_meshlet*>(prim.meshlet_buffer)[dr.meshlet_index];
        const Mat4 mvp = mat4_mul(a.vp, load_mat4(a.s.transforms + 16ull * dr.transform_index));
        const uint8_t* tri = reinterpret_cast<const uint8_t*>(prim.primitive_index_buffer) + ml.triangle_offset + 3u * t;
        const uint32_t* vidx = reinterpret_cast<const uint32_t*>(prim.vertex_index_buffer) + ml.vertex_offset;
        const vkgv_vertex* verts = reinterpret_cast<const vkgv_vertex*>(prim.vertex_buffer);
        float3 hv[3];
        float col[3][4];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const vkgv_vertex& vx = verts[vidx[tri[k]]];
            const float4 cl = mat4_point(mvp, vx.position[0], vx.position[1], vx.position[2]);
            hv[k] = make_float3(cl.x, cl.y, cl.w);
            const uint32_t packed = *reinterpret_cast<const uint32_t*>(vx.color);
#pragma unroll
            for (int j = 0; j < 4; ++j) col[k][j] = (float)((packed >> (8 * j)) & 0xFFu) / 255.0f;
        }
        const float hw = (float)a.W * 0.5f, hh = (float)a.H * 0.5f;
        const float xn = (((float)px + 0.5f) - hw) / hw;
        const float yn = (((float)py + 0.5f) - hh) / hh;
        const float3 r0 = cross3(hv[1], hv[2]), r1 = cross3(hv[2], hv[0]), r2 = cross3(hv[0], hv[1]);
        float l0 = (r0.x * xn + r0.y * yn) + r0.z;
        float l1 = (r1.x * xn + r1.y * yn) + r1.z;
        float l2 = (r2.x * xn + r2.y * yn) + r2.z;
        const float sum = (l0 + l1) + l2;
        l0 = l0 / sum; l1 = l1 / sum; l2 = l2 / sum;
#pragma unroll
        for (int j = 0; j < 4; ++j) c[j] = c[j] * ((l0 * col[0][j] + l1 * col[1][j]) + l2 * col[2][j]);
    }
    const uint32_t r8 = unorm8(from_linear(c[0])), g8 = unorm8(from_linear(c[1])), b8 = unorm8(from_linear(c[2]));
    const uint32_t a8 = unorm8(c[3]);
    a.color[i] = r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
}

} // namespace

void launch_resolve(const SceneDev& s, const ViewDev& v, const Mat4& vp, uint32_t flags, hipStream_t st)
{
    ResolveArgs a;
    a.s = s;
    a.vp = vp;
    a.keys = v.keys;
    a.color = v.color;
    a.W = v.width;
    a.H = v.height;
    a.flags = flags;
    const dim3 grid((v.width + 63) / 64, (v.height + 3) / 4);
    resolve_kernel<<<grid, 256, 0, st>>>(a);
}

} // namespace vkgv
