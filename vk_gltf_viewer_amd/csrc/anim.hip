// Per-frame transform update on the device (World::updateDrawBuffers,
// world.cpp:346-354): animation sampling per node (AnimationSampler::sample,
// assets.hpp:52-110; channel application, world.cpp:192-216), then the world
// matrix of every mesh node down its ancestor chain (World::iterateNode,
// world.cpp:187-228: parent * T * R * S).
//
// Kernel 1, one thread per node: the node's static TRS, overwritten by each
// of its channels in order, to a per-node TRS buffer.  Kernel 2, one thread
// per mesh node: the product along its root-to-node chain, in the same
// operation order as the host traversal (so an unanimated hierarchy gives
// the host's matrices bit for bit), written to the transform buffer.  A
// chain is at most 64 nodes; the work per mesh node is 3 * depth matrix
// products, with no dependence between threads.
//
// slerp evaluates acos and sin in double precision with fixed series (below)
// instead of a math library, so that every platform computes the same bits.
#include "internal.hpp"

namespace vkgv {
namespace {

// asin(y), |y| <= 1/2: Maclaurin series y * sum c_n y^(2n) with
// c_{n+1} / c_n = (2n+1)^2 / ((2n+2)(2n+3)), in Horner form (31 terms).
__device__ double series_asin(double y)
{
    const double y2 = y * y;
    double acc = 1.0;
    for (int n = 29; n >= 0; --n)
        acc = 1.0 + (y2 * acc) * ((double)((2 * n + 1) * (2 * n + 1)) / (double)((2 * n + 2) * (2 * n + 3)));
    return y * acc;
}

// acos(x), 0 <= x <= 1.
__device__ double series_acos(double x)
{
    if (x <= 0.5) return 1.5707963267948966 - series_asin(x);
    return 2.0 * series_asin(sqrt((1.0 - x) * 0.5));
}

// sin(a), 0 <= a <= pi/2: Taylor series, 16 terms, in Horner form.
__device__ double series_sin(double a)
{
    const double a2 = a * a;
    double acc = 1.0;
    for (int n = 14; n >= 0; --n) acc = 1.0 - (a2 * acc) / (double)((2 * n + 2) * (2 * n + 3));
    return a * acc;
}

// Output k of the sampler's values (N floats per output).
template <int N>
__device__ __forceinline__ void load_out(const float* v, uint32_t k, float* o)
{
#pragma unroll
    for (int i = 0; i < N; ++i) o[i] = v[(size_t)k * N + i];
}

// AnimationSampler::sample<path>(time) (assets.hpp:52-110); N = 3 for
// translation and scale, 4 for rotation.
template <int N>
__device__ __forceinline__ void sample(const AnimDev& d, const vkgv_anim_sampler& sp, float time, float* out)
{
    const float* in = d.times + sp.input_offset;
    const float* val = d.values + sp.output_offset;
    const uint32_t cnt = sp.input_count;
    const bool cubic = sp.interpolation == VKGV_INTERP_CUBICSPLINE;
    time = fmodf(time, in[cnt - 1]);                        // loop the animation
    // std::lower_bound (first key >= time), with libstdc++'s probe sequence so
    // that even unsorted times pick the key the reference picks
    uint32_t lo = 0, len = cnt;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (in[lo + half] < time) { lo = lo + half + 1; len = len - half - 1; }
        else len = half;
    }
    if (lo == 0) { load_out<N>(val, cubic ? 1u : 0u, out); return; }
    if (lo == cnt) { load_out<N>(val, cubic ? 3u * cnt - 2u : cnt - 1u, out); return; }
    const uint32_t i = lo;
    const float t = (time - in[i - 1]) / (in[i] - in[i - 1]);
    if (sp.interpolation == VKGV_INTERP_STEP) { load_out<N>(val, i - 1, out); return; }
    if (!cubic) {
        float a[N], b[N];
        load_out<N>(val, i - 1, a);
        load_out<N>(val, i, b);
        if (N == 3) {
#pragma unroll
            for (int k = 0; k < N; ++k) out[k] = a[k] * (1.0f - t) + b[k] * t;
            return;
        }
        // slerp (shortest arc; linear when the keys nearly coincide)
        float cosv = 0.0f;
#pragma unroll
        for (int k = 0; k < N; ++k) cosv = k == 0 ? a[0] * b[0] : cosv + a[k] * b[k];
        if (cosv < 0.0f) {
#pragma unroll
            for (int k = 0; k < N; ++k) b[k] = -b[k];
            cosv = -cosv;
        }
        if (cosv > 1.0f - 1.1920928955078125e-7f) {
#pragma unroll
            for (int k = 0; k < N; ++k) out[k] = a[k] * (1.0f - t) + b[k] * t;
            return;
        }
        const double ang = series_acos((double)cosv);
        const double sa = series_sin(ang);
        const double w0 = series_sin((1.0 - (double)t) * ang) / sa;
        const double w1 = series_sin((double)t * ang) / sa;
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = (float)(w0 * (double)a[k] + w1 * (double)b[k]);
        return;
    }
    // cubic Hermite spline (glTF 2.0 appendix C): in-tangent, value, out-tangent per key
    const float t2 = t * t, t3 = t2 * t;
    const float dt = in[i] - in[i - 1];
    float v0[N], b0[N], v1[N], a1[N];
    load_out<N>(val, 3u * (i - 1) + 1u, v0);
    load_out<N>(val, 3u * (i - 1) + 2u, b0);
    load_out<N>(val, 3u * i + 1u, v1);
    load_out<N>(val, 3u * i + 0u, a1);
    const float h00 = (2.0f * t3 - 3.0f * t2) + 1.0f;
    const float h10 = (t3 - 2.0f * t2) + t;
    const float h01 = -2.0f * t3 + 3.0f * t2;
    const float h11 = t3 - t2;
#pragma unroll
    for (int k = 0; k < N; ++k) out[k] = ((v0[k] * h00 + (b0[k] * h10) * dt) + v1[k] * h01) + (a1[k] * h11) * dt;
    if (N == 4) {
        float l2 = 0.0f;
#pragma unroll
        for (int k = 0; k < N; ++k) l2 = k == 0 ? out[0] * out[0] : l2 + out[k] * out[k];
        const float l = sqrtf(l2);
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = out[k] / l;
    }
}

__global__ __launch_bounds__(256) void anim_nodes_kernel(AnimDev d, float time)
{
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= d.node_count) return;
    const vkgv_node& nd = d.nodes[n];
    NodeTrs x;
#pragma unroll
    for (int k = 0; k < 3; ++k) { x.t[k] = nd.translation[k]; x.s[k] = nd.scale[k]; }
#pragma unroll
    for (int k = 0; k < 4; ++k) x.q[k] = nd.rotation[k];
    const uint32_t first = d.chan_first[n], count = d.chan_first[n + 1] - first;
    for (uint32_t c = 0; c < count; ++c) {
        const vkgv_anim_channel ch = d.channels[first + c];
        const vkgv_anim_sampler sp = d.samplers[ch.sampler];
        if (ch.path == VKGV_PATH_ROTATION) {
            sample<4>(d, sp, time, x.q);
        } else {
            float v[3];
            sample<3>(d, sp, time, v);
            if (ch.path == VKGV_PATH_TRANSLATION) { x.t[0] = v[0]; x.t[1] = v[1]; x.t[2] = v[2]; }
            else { x.s[0] = v[0]; x.s[1] = v[1]; x.s[2] = v[2]; }
        }
    }
    d.trs[n] = x;
}

// parent * T * R * S with the host traversal's operation order.
__device__ __forceinline__ Mat4 apply_trs(const Mat4& parent, const NodeTrs& n)
{
    const Mat4 T = {{ 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, n.t[0], n.t[1], n.t[2], 1 }};
    const float x = n.q[0], y = n.q[1], z = n.q[2], w = n.q[3];
    const Mat4 R = {{ 1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y), 0.0f,
                      2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x), 0.0f,
                      2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y), 0.0f,
                      0.0f, 0.0f, 0.0f, 1.0f }};
    const Mat4 S = {{ n.s[0], 0, 0, 0, 0, n.s[1], 0, 0, 0, 0, n.s[2], 0, 0, 0, 0, 1 }};
    return mat4_mul(mat4_mul(mat4_mul(parent, T), R), S);
}

__global__ __launch_bounds__(256) void world_transforms_kernel(AnimDev d, float* out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= d.mesh_count) return;
    Mat4 m = {{ 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 }};
    const uint32_t first = d.chain_first[i], len = d.chain_first[i + 1] - first;
    for (uint32_t k = 0; k < len; ++k) m = apply_trs(m, d.trs[d.chain_nodes[first + k]]);
    float4* o = reinterpret_cast<float4*>(out + 16ull * i);
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = make_float4(m.m[4 * c], m.m[4 * c + 1], m.m[4 * c + 2], m.m[4 * c + 3]);
}

} // namespace

void launch_update_transforms(const AnimDev& d, float time, float* transforms, hipStream_t st)
{
    if (d.node_count) anim_nodes_kernel<<<(d.node_count + 255) / 256, 256, 0, st>>>(d, time);
    if (d.mesh_count) world_transforms_kernel<<<(d.mesh_count + 255) / 256, 256, 0, st>>>(d, transforms);
}

} // namespace vkgv
