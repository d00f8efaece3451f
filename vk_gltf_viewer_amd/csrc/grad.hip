// Differentiable frame: the antialiased linear image of a view's last frame
// and its gradient with respect to the instance transforms, computed from the
// frame's own 64-bit visibility buffer (no second raster).
//
// Image: each pixel's material albedo factor (the flat resolve before the
// sRGB encode, visbuffer_resolve.comp.glsl:17-41; background 0), with the
// analytic edge antialiasing of the differentiable form: for every
// horizontally or vertically adjacent pixel pair with different colours
// (different materials, or one side background), the front pixel (the
// non-background one, else the one with the smaller key, i.e. nearer) keeps
// its triangle; the segment from its centre to the other centre leaves the
// triangle through the edge whose edge function changes sign first, at t in
// [0, 1]; t < 1/2 blends (1/2 - t) of the back colour into the front pixel,
// else (t - 1/2) of the front colour into the back pixel.  Pairs of equal
// colour change nothing and are skipped.
//
// Geometry: id = draw << 7 | triangle -> the draw's raster record (vertex,
// triangle and vertex-index offsets into the packed mesh buffers, transform
// index, material) -> the triangle's object-space positions p; the clip
// position VP M (p, 1) is formed in double (VP, M are the camera's and the
// instance's float matrices), and so are the screen offsets from the front
// pixel's centre: the edge term scales as 1 / (edge function difference)^2,
// which float32 clip positions do not resolve.
//
// A first pass writes each pixel's material index, so that the pairs of equal
// colour (nearly all of them) are skipped after two coalesced loads.
//
// Backward: one thread per pair; d loss / d M_tr = sum over the crossed
// edge's two vertices of (VP^T g_clip) (p, 1)^T.  Lanes of a wave mostly
// share a transform, so each wave reduces its 16-value contributions per
// distinct transform with lane shuffles before one lane adds them.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace vkgv {
namespace {

constexpr uint32_t kGradBlock = 256;

struct GradArgs {
    SceneDev s;
    const unsigned long long* keys;
    const uint32_t* mat;    // per pixel: material index, kNoMaterial on background
    uint32_t W, H;
    double vp[16];          // column-major
};

constexpr uint32_t kNoMaterial = 0xFFFFFFFFu;

__device__ __forceinline__ bool key_id(unsigned long long k, uint32_t& id)
{
    if (k == ~0ull) return false;
    id = 0xFFFFFFFEu - (uint32_t)k;
    return true;
}

// material of a key's draw; kNoMaterial on background or an id outside the draw list
__device__ __forceinline__ uint32_t material_of_key(const SceneDev& s, unsigned long long k)
{
    uint32_t id;
    if (!key_id(k, id)) return kNoMaterial;
    const uint32_t d = id >> VKGV_TRIANGLE_BITS;
    if (d >= s.draw_count) return kNoMaterial;
    return s.rrec[d].material & 0x7FFFFFFFu;
}

__device__ __forceinline__ float3 albedo(const SceneDev& s, uint32_t m)
{
    if (m == kNoMaterial) return make_float3(0.0f, 0.0f, 0.0f);
    const vkgv_material& mt = s.materials[m];
    return make_float3(mt.albedo_factor[0], mt.albedo_factor[1], mt.albedo_factor[2]);
}

struct DPair {
    size_t f, b;
    bool to_front;
    uint32_t tr;
    double alpha, ef, eb;
    double sa[2], sb[2], cb[2];
    double pa[4], pb[4];
    double oa[3], ob[3];
};

__device__ __forceinline__ double efn(const double* a, const double* b, const double* p)
{
    return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]);
}

// Pair (px, py) - right (dir 0) or lower (dir 1) neighbour.
__device__ bool pair_eval(const GradArgs& a, int px, int py, int dir, DPair& e)
{
    const int qx = px + (dir == 0), qy = py + (dir == 1);
    if (qx >= (int)a.W || qy >= (int)a.H) return false;
    const size_t ip = (size_t)py * a.W + px, iq = (size_t)qy * a.W + qx;
    const uint32_t mp = a.mat[ip], mq = a.mat[iq];
    if (mp == mq) return false;                        // equal colours (or both background): no blend
    const bool hp = mp != kNoMaterial, hq = mq != kNoMaterial;
    const unsigned long long kp = a.keys[ip], kq = a.keys[iq];
    const bool p_front = !hp ? false : (!hq ? true : kp <= kq);
    e.f = p_front ? ip : iq;
    e.b = p_front ? iq : ip;
    uint32_t id;
    key_id(p_front ? kp : kq, id);
    const int fx = p_front ? px : qx, fy = p_front ? py : qy;
    const double step = p_front ? 1.0 : -1.0;
    e.cb[0] = dir == 0 ? step : 0.0;
    e.cb[1] = dir == 0 ? 0.0 : step;
    const RasterRec r = a.s.rrec[id >> VKGV_TRIANGLE_BITS];
    const uint32_t t = id & ((1u << VKGV_TRIANGLE_BITS) - 1u);
    if (t >= ((r.counts >> 8) & 0xFFu)) return false;
    e.tr = r.transform;
    const float* m = a.s.transforms + 16ull * r.transform;
    double M[16];
    for (int i = 0; i < 16; ++i) M[i] = (double)m[i];
    const double hw = a.W * 0.5, hh = a.H * 0.5;
    const double xc = (fx + 0.5) / hw - 1.0, yc = (fy + 0.5) / hh - 1.0;
    double P[3][4], S[3][2], O[3][3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t li = a.s.tri_base[r.tri + 3u * t + (uint32_t)k];
        const uint32_t vi = a.s.vidx_base[r.vidx + li];
        const float* pp = reinterpret_cast<const float*>(reinterpret_cast<const uint8_t*>(a.s.vtx_base) + r.vtx) +
                          (size_t)vi * 6;
        O[k][0] = pp[0]; O[k][1] = pp[1]; O[k][2] = pp[2];
        double w4[4];
        for (int rr = 0; rr < 4; ++rr) w4[rr] = M[rr] * O[k][0] + M[4 + rr] * O[k][1] + M[8 + rr] * O[k][2] + M[12 + rr];
        for (int rr = 0; rr < 4; ++rr)
            P[k][rr] = a.vp[rr] * w4[0] + a.vp[4 + rr] * w4[1] + a.vp[8 + rr] * w4[2] + a.vp[12 + rr] * w4[3];
        if (!(P[k][3] > 0.0)) return false;
        S[k][0] = hw * ((P[k][0] - xc * P[k][3]) / P[k][3]);
        S[k][1] = hh * ((P[k][1] - yc * P[k][3]) / P[k][3]);
    }
    const double area = efn(S[0], S[1], S[2]);
    if (area == 0.0) return false;
    const double sgn = area > 0.0 ? 1.0 : -1.0;
    const double zero[2] = { 0.0, 0.0 };
    double best = 2.0, bef = 0.0, beb = 0.0;
    int kb = -1;
    for (int k = 0; k < 3; ++k) {
        const double ef = sgn * efn(S[k], S[(k + 1) % 3], zero);
        const double eb = sgn * efn(S[k], S[(k + 1) % 3], e.cb);
        if (!(ef >= 0.0 && eb <= 0.0 && ef > eb)) continue;
        const double tt = ef / (ef - eb);
        if (tt < best) { best = tt; kb = k; bef = ef; beb = eb; }
    }
    if (kb < 0) return false;
    e.to_front = best < 0.5;
    e.alpha = e.to_front ? 0.5 - best : best - 0.5;
    const int ka = kb, kc = (kb + 1) % 3;
    for (int j = 0; j < 4; ++j) { e.pa[j] = P[ka][j]; e.pb[j] = P[kc][j]; }
    for (int j = 0; j < 3; ++j) { e.oa[j] = O[ka][j]; e.ob[j] = O[kc][j]; }
    for (int j = 0; j < 2; ++j) { e.sa[j] = S[ka][j]; e.sb[j] = S[kc][j]; }
    e.ef = sgn * bef;
    e.eb = sgn * beb;
    return true;
}

// Per pixel: the material of its draw (kNoMaterial on background).
__global__ __launch_bounds__(kGradBlock) void pixel_material_kernel(SceneDev s, const unsigned long long* __restrict__ keys,
                                                                   size_t n, uint32_t* __restrict__ mat)
{
    const size_t p = (size_t)blockIdx.x * kGradBlock + threadIdx.x;
    if (p < n) mat[p] = material_of_key(s, keys[p]);
}

// Gather: pixel p = its colour + the blends of its four pairs that target it.
__global__ __launch_bounds__(kGradBlock) void aa_image_kernel(GradArgs a, float* __restrict__ image)
{
    const size_t p = (size_t)blockIdx.x * kGradBlock + threadIdx.x;
    if (p >= (size_t)a.W * a.H) return;
    const int px = (int)(p % a.W), py = (int)(p / a.W);
    const float3 c = albedo(a.s, a.mat[p]);
    double acc[3] = { c.x, c.y, c.z };
    const int sx[4] = { px - 1, px, px, px }, sy[4] = { py, py, py - 1, py }, dr[4] = { 0, 0, 1, 1 };
    for (int q = 0; q < 4; ++q) {
        if (sx[q] < 0 || sy[q] < 0) continue;
        DPair e;
        if (!pair_eval(a, sx[q], sy[q], dr[q], e)) continue;
        const size_t tgt = e.to_front ? e.f : e.b, src = e.to_front ? e.b : e.f;
        if (tgt != p) continue;
        const float3 cs = albedo(a.s, a.mat[src]);
        acc[0] += e.alpha * ((double)cs.x - c.x);
        acc[1] += e.alpha * ((double)cs.y - c.y);
        acc[2] += e.alpha * ((double)cs.z - c.z);
    }
    image[3 * p] = (float)acc[0];
    image[3 * p + 1] = (float)acc[1];
    image[3 * p + 2] = (float)acc[2];
}

__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// One thread per pair (dir-major: threads [0, W*H) horizontal pairs, then vertical).
__global__ __launch_bounds__(kGradBlock) void aa_backward_kernel(GradArgs a, const float* __restrict__ g_image,
                                                                 double* __restrict__ g_transforms)
{
    const size_t npx = (size_t)a.W * a.H;
    const size_t g = (size_t)blockIdx.x * kGradBlock + threadIdx.x;
    double G[16];
    int key = -1;
    if (g < 2 * npx) {
        const int dir = (int)(g / npx);
        const size_t p = g % npx;
        DPair e;
        if (pair_eval(a, (int)(p % a.W), (int)(p / a.W), dir, e)) {
            const size_t tgt = e.to_front ? e.f : e.b, src = e.to_front ? e.b : e.f;
            const float3 ct = albedo(a.s, a.mat[tgt]), cs = albedo(a.s, a.mat[src]);
            const double galpha = (double)g_image[3 * tgt] * ((double)cs.x - ct.x) +
                                  (double)g_image[3 * tgt + 1] * ((double)cs.y - ct.y) +
                                  (double)g_image[3 * tgt + 2] * ((double)cs.z - ct.z);
            const double gt = e.to_front ? -galpha : galpha;
            const double den = e.ef - e.eb;
            const double gef = gt * (-e.eb / (den * den)), geb = gt * (e.ef / (den * den));
            // e(P) = (bx - ax)(Py - ay) - (by - ay)(Px - ax) at P = 0 (front) and cb (back)
            const double gax = gef * (e.sb[1] - 0.0) + geb * (e.sb[1] - e.cb[1]);
            const double gay = gef * (0.0 - e.sb[0]) + geb * (e.cb[0] - e.sb[0]);
            const double gbx = gef * (0.0 - e.sa[1]) + geb * (e.cb[1] - e.sa[1]);
            const double gby = gef * (e.sa[0] - 0.0) + geb * (e.sa[0] - e.cb[0]);
            const double hw = a.W * 0.5, hh = a.H * 0.5;
            for (int k = 0; k < 16; ++k) G[k] = 0.0;
            const double* pc[2] = { e.pa, e.pb };
            const double* po[2] = { e.oa, e.ob };
            const double gsx[2] = { gax, gbx }, gsy[2] = { gay, gby };
            for (int v = 0; v < 2; ++v) {
                const double* c = pc[v];
                const double gc0 = gsx[v] * hw / c[3], gc1 = gsy[v] * hh / c[3];
                const double gc3 = -(gsx[v] * hw * c[0] + gsy[v] * hh * c[1]) / (c[3] * c[3]);
                double gw[4];   // VP^T g_clip (g_clip.z = 0)
                for (int j = 0; j < 4; ++j) gw[j] = a.vp[j * 4] * gc0 + a.vp[j * 4 + 1] * gc1 + a.vp[j * 4 + 3] * gc3;
                const double ph[4] = { po[v][0], po[v][1], po[v][2], 1.0 };
                for (int col = 0; col < 4; ++col)
                    for (int r = 0; r < 4; ++r) G[col * 4 + r] += gw[r] * ph[col];
            }
            key = (int)e.tr;
        }
    }
    // wave reduction per distinct transform; every lane takes part in every shuffle
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(key >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int k0 = __shfl(key, leader, 64);
        const bool mine = key == k0;
        for (int j = 0; j < 16; ++j) {
            const double s = wave_sum(mine ? G[j] : 0.0);
            if (lane == leader) atomicAdd(&g_transforms[16ull * (uint32_t)k0 + j], s);
        }
        pending &= ~__ballot(mine);
    }
}

inline unsigned grad_blocks(size_t n) { return (unsigned)((n + kGradBlock - 1) / kGradBlock); }

GradArgs make_args(const SceneDev& s, const unsigned long long* keys, const uint32_t* mat, uint32_t W, uint32_t H,
                   const float* vp)
{
    GradArgs a;
    a.s = s;
    a.keys = keys;
    a.mat = mat;
    a.W = W;
    a.H = H;
    for (int i = 0; i < 16; ++i) a.vp[i] = (double)vp[i];
    return a;
}

} // namespace

void launch_aa_image(const SceneDev& s, const unsigned long long* keys, uint32_t* mat, uint32_t W, uint32_t H,
                     const float* vp, float* image, hipStream_t st)
{
    const size_t n = (size_t)W * H;
    if (n == 0) return;
    pixel_material_kernel<<<grad_blocks(n), kGradBlock, 0, st>>>(s, keys, n, mat);
    aa_image_kernel<<<grad_blocks(n), kGradBlock, 0, st>>>(make_args(s, keys, mat, W, H, vp), image);
}

void launch_aa_backward(const SceneDev& s, const unsigned long long* keys, uint32_t* mat, uint32_t W, uint32_t H,
                        const float* vp, const float* g_image, double* g_transforms, hipStream_t st)
{
    const size_t n = (size_t)W * H;
    if (n == 0) return;
    pixel_material_kernel<<<grad_blocks(n), kGradBlock, 0, st>>>(s, keys, n, mat);
    aa_backward_kernel<<<grad_blocks(2 * n), kGradBlock, 0, st>>>(make_args(s, keys, mat, W, H, vp), g_image,
                                                                  g_transforms);
}

} // namespace vkgv
