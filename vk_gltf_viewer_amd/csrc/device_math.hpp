// Device-side math of the frame: GLSL semantics of the reference shaders with
// a fixed floating-point evaluation order.  The library is compiled with
// -ffp-contract=off, so every expression below rounds exactly as written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vkgv.h"

namespace vkgv {

// GLSL min/max/clamp (GLSL 4.60 section 8.3: min(x, y) = y < x ? y : x).
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

struct Mat4 { float m[16]; };   // column-major, m[c*4 + r]

__device__ __forceinline__ Mat4 load_mat4(const float* p)
{
    Mat4 r;
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 v = q[i];
        r.m[4 * i] = v.x; r.m[4 * i + 1] = v.y; r.m[4 * i + 2] = v.z; r.m[4 * i + 3] = v.w;
    }
    return r;
}

// A * B, each element ((a0*b0 + a1*b1) + a2*b2) + a3*b3.
__device__ __forceinline__ Mat4 mat4_mul(const Mat4& a, const Mat4& b)
{
    Mat4 t;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            t.m[c * 4 + r] = ((a.m[r] * b.m[c * 4] + a.m[4 + r] * b.m[c * 4 + 1]) + a.m[8 + r] * b.m[c * 4 + 2])
                             + a.m[12 + r] * b.m[c * 4 + 3];
    return t;
}

// m * (x, y, z, 1); the last term multiplies by exactly 1.
__device__ __forceinline__ float4 mat4_point(const Mat4& m, float x, float y, float z)
{
    float4 o;
    o.x = ((m.m[0] * x + m.m[4] * y) + m.m[8] * z) + m.m[12] * 1.0f;
    o.y = ((m.m[1] * x + m.m[5] * y) + m.m[9] * z) + m.m[13] * 1.0f;
    o.z = ((m.m[2] * x + m.m[6] * y) + m.m[10] * z) + m.m[14] * 1.0f;
    o.w = ((m.m[3] * x + m.m[7] * y) + m.m[11] * z) + m.m[15] * 1.0f;
    return o;
}

// Sign-relevant determinant of the instance transform (visbuffer.mesh.glsl:71):
// expansion down column 0 with the 2x2 minors of columns 2 and 3.
__device__ __forceinline__ float det4(const Mat4& a)
{
    const float* m = a.m;
#define VKGV_T(ra, rb) (m[8 + (ra)] * m[12 + (rb)] - m[8 + (rb)] * m[12 + (ra)])
    const float t23 = VKGV_T(2, 3), t13 = VKGV_T(1, 3), t12 = VKGV_T(1, 2);
    const float t03 = VKGV_T(0, 3), t02 = VKGV_T(0, 2), t01 = VKGV_T(0, 1);
#undef VKGV_T
    const float c0 = (m[5] * t23 - m[6] * t13) + m[7] * t12;
    const float c1 = (m[4] * t23 - m[6] * t03) + m[7] * t02;
    const float c2 = (m[4] * t13 - m[5] * t03) + m[7] * t01;
    const float c3 = (m[4] * t12 - m[5] * t02) + m[6] * t01;
    return ((m[0] * c0 - m[1] * c1) + m[2] * c2) - m[3] * c3;
}

// determinant(mat3(a.xyw, b.xyw, c.xyw)) (visbuffer.mesh.glsl:90-93).
__device__ __forceinline__ float det3_xyw(float4 a, float4 b, float4 c)
{
    return (a.x * (b.y * c.w - c.y * b.w) - b.x * (a.y * c.w - c.y * a.w)) + c.x * (a.y * b.w - b.y * a.w);
}

// ---------------------------------------------------------------------------
// Visibility-buffer key: (~float_bits(depth) << 32) | (0xFFFFFFFE - id).
__device__ __forceinline__ uint64_t make_key(float depth, uint32_t id)
{
    return ((uint64_t)(~__float_as_uint(depth + 0.0f)) << 32) | (uint64_t)(0xFFFFFFFEu - id);
}
__device__ __forceinline__ float key_depth(uint64_t key) { return __uint_as_float(~(uint32_t)(key >> 32)); }
__device__ __forceinline__ uint32_t key_id(uint64_t key) { return 0xFFFFFFFEu - (uint32_t)key; }

// ---------------------------------------------------------------------------
// Hi-Z pyramid geometry passed to kernels.
struct PyramidDesc {
    uint32_t levels;
    uint32_t w[16], h[16];
    uint32_t off[16];      // float offsets into the pyramid buffer
};

// Linear-filter footprint along one axis for unnormalized coordinate u:
// texel floor(u), and floor(u)+1 only when its weight frac(u) is non-zero
// (the MIN reduction takes the texels with non-zero weight), clamp-to-edge.
__device__ __forceinline__ void axis_footprint(float u, int size, int& t0, int& t1)
{
    const float fl = floorf(u);
    const float a = u - fl;
    const int i0 = (int)fl;
    t0 = min(max(i0, 0), size - 1);
    t1 = a != 0.0f ? min(max(i0 + 1, 0), size - 1) : t0;
}

// floor(log2(x)) exactly from the binary exponent: 0 -> -inf, NaN/neg -> NaN.
__device__ __forceinline__ float floor_log2(float x)
{
    if (!(x >= 0.0f)) return __int_as_float(0x7fc00000);
    if (x == 0.0f) return -__int_as_float(0x7f800000);
    const uint32_t b = __float_as_uint(x);
    const int e = (int)(b >> 23);
    if (e == 255) return __int_as_float(0x7f800000);
    if (e == 0) return (float)(-127 - (int)__clz(b << 9));   // subnormal: exponent of the leading bit
    return (float)(e - 127);
}

} // namespace vkgv
