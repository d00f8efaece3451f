// Hi-Z depth pyramid (hiz_reduce.comp.glsl:21-31, application.cpp:435-529, 951-1003).
//
// Level k (0-based) is max(1, W >> (k+1)) x max(1, H >> (k+1)); texel p of a
// level samples its source (the depth image for level 0, level k-1 after) at
// (p + 0.5) / dstSize through the MIN-reduction linear sampler, i.e. the
// minimum of the footprint texels with non-zero bilinear weight.
//
// Two launches instead of one per level:
//  1. hiz_tiles: one 256-thread block per 32x32 texel tile of level 0, read
//     straight from the 64-bit visibility buffer; the following levels are
//     reduced in LDS for as long as every source level is exactly twice its
//     destination (then the footprint of texel p is {2p, 2p+1} and stays
//     inside the tile), writing each level once;
//  2. hiz_tail: one 1024-thread block per view for the remaining (small)
//     levels with the general footprint.
// Both cover every view of a batch in one launch.
#include "internal.hpp"

namespace vkgv {
namespace {

__device__ __forceinline__ void footprint(uint32_t p, uint32_t dst, uint32_t src, int& t0, int& t1)
{
    const float u = (((float)p + 0.5f) / (float)dst) * (float)src - 0.5f;
    axis_footprint(u, (int)src, t0, t1);
}

// One block per 32x32 tile of a view's level 0 (block map entry = view |
// tile << 8).
__global__ __launch_bounds__(256) void hiz_tiles_kernel(BatchDev b)
{
    __shared__ float s_a[32 * 32];
    __shared__ float s_b[16 * 16];
    const uint32_t e = b.map_hiz[blockIdx.x];
    const ViewDev& v = batch_view(b, e & 0xFFu);
    const uint32_t tile = e >> 8;
    const unsigned long long* keys = v.keys;
    const uint32_t W = v.width, H = v.height;
    const PyramidDesc& pd = v.pyr;
    const uint32_t tile_levels = v.hiz_tile_levels;
    const uint32_t w0 = pd.w[0], h0 = pd.h[0];
    const uint32_t bx = (tile % v.hiz_tiles_x) * 32u, by = (tile / v.hiz_tiles_x) * 32u;
    float* lvl0 = v.pyramid + pd.off[0];
    for (uint32_t t = threadIdx.x; t < 1024u; t += 256u) {
        const uint32_t x = bx + (t & 31u), y = by + (t >> 5);
        float d = 0.0f;
        if (x < w0 && y < h0) {
            int x0, x1, y0, y1;
            footprint(x, w0, W, x0, x1);
            footprint(y, h0, H, y0, y1);
            const unsigned long long* r0 = keys + (size_t)y0 * W;
            const unsigned long long* r1 = keys + (size_t)y1 * W;
            d = key_depth(r0[x0]);
            d = gmin(d, key_depth(r0[x1]));
            d = gmin(d, key_depth(r1[x0]));
            d = gmin(d, key_depth(r1[x1]));
            lvl0[(size_t)y * w0 + x] = d;
        }
        s_a[t] = d;
    }
    float* src = s_a;
    float* dst = s_b;
    uint32_t ts = 32u;   // tile size of the source level in LDS
    for (uint32_t k = 1; k < tile_levels; ++k) {
        __syncthreads();
        const uint32_t nts = ts >> 1;
        const uint32_t wk = pd.w[k], hk = pd.h[k];
        const uint32_t ox = bx >> k, oy = by >> k;
        float* out = v.pyramid + pd.off[k];
        for (uint32_t t = threadIdx.x; t < nts * nts; t += 256u) {
            const uint32_t i = t % nts, j = t / nts;
            const uint32_t x = ox + i, y = oy + j;
            const float* s0 = src + (2u * j) * ts + 2u * i;
            // footprint {2i, 2i+1} x {2j, 2j+1}, taken in the order (x0,y0),(x1,y0),(x0,y1),(x1,y1)
            float d = s0[0];
            d = gmin(d, s0[1]);
            d = gmin(d, s0[ts]);
            d = gmin(d, s0[ts + 1]);
            if (x < wk && y < hk) out[(size_t)y * wk + x] = d;
            dst[t] = d;
        }
        float* tmp = src; src = dst; dst = tmp;
        ts = nts;
    }
}

// The levels after the tile pass's, one block per view.  A level small
// enough for LDS is also kept there, so that the following levels read LDS
// instead of global memory (one global round trip for the whole tail when
// its first level fits).  A texel's footprint depends only on its column
// (x0, x1) and row (y0, y1), so each level's column and row footprints are
// computed once into LDS (dw + dh <= kTailA + 1 entries) instead of per texel.
constexpr uint32_t kTailA = kHizTailTexels, kTailB = kHizTailTexels / 2;   // floats of the two LDS level buffers
constexpr uint32_t kTailPer = kTailA / 1024u;                             // texels per thread of a tail level
static_assert(kTailA % 1024u == 0, "the tail's levels are whole rounds of its 1024 threads");

__global__ __launch_bounds__(1024) void hiz_tail_kernel(BatchDev b)
{
    __shared__ float s_a[kTailA];
    __shared__ float s_b[kTailB];
    __shared__ uint32_t s_fp[kTailA + 1];   // footprints: columns [0, dw), then rows [dw, dw + dh): t0 | t1 << 16
    const ViewDev& v = batch_view(b, blockIdx.x);
    const PyramidDesc& pd = v.pyr;
    float* pyr = v.pyramid;
    const float* lsrc = nullptr;          // LDS copy of level k-1, if any
    bool next_a = true;                   // LDS buffer for the next level kept in LDS
    for (uint32_t k = v.hiz_tail_start; k < pd.levels; ++k) {
        const uint32_t sw = pd.w[k - 1], sh = pd.h[k - 1], dw = pd.w[k], dh = pd.h[k];
        const float* src = lsrc ? lsrc : pyr + pd.off[k - 1];
        float* dst = pyr + pd.off[k];
        float* ldst = nullptr;
        if (dw * dh <= (next_a ? kTailA : kTailB)) ldst = next_a ? s_a : s_b;
        for (uint32_t i = threadIdx.x; i < dw + dh; i += 1024u) {
            int t0, t1;
            if (i < dw) footprint(i, dw, sw, t0, t1);
            else footprint(i - dw, dh, sh, t0, t1);
            s_fp[i] = (uint32_t)t0 | ((uint32_t)t1 << 16);
        }
        __syncthreads();
        // a level holds at most kTailA texels: kTailPer per thread, all of a
        // thread's loads issued before the first minimum (the first level is
        // read from global memory, where one round trip each would add up)
        const uint32_t n = dw * dh;
#pragma unroll
        for (uint32_t r = 0; r < kTailPer; ++r) {
            const uint32_t t = threadIdx.x + r * 1024u;
            if (t < n) {
                const uint32_t y = t / dw, x = t - y * dw;
                const uint32_t fx = s_fp[x], fy = s_fp[dw + y];
                const uint32_t x0 = fx & 0xFFFFu, x1 = fx >> 16, y0 = fy & 0xFFFFu, y1 = fy >> 16;
                float d = src[y0 * sw + x0];
                d = gmin(d, src[y0 * sw + x1]);
                d = gmin(d, src[y1 * sw + x0]);
                d = gmin(d, src[y1 * sw + x1]);
                dst[t] = d;
                if (ldst) ldst[t] = d;
            }
        }
        __syncthreads();
        lsrc = ldst;
        if (ldst) next_a = !next_a;
    }
}

// Levels after the tile pass that are too large for the tail's single block
// (a source level of odd size ends the tile pass early, e.g. at level 1 of a
// view of odd width): one launch per level for every view, 256 texels per
// block; block (x, view).  Same footprint as the tail.
__global__ __launch_bounds__(256) void hiz_level_kernel(BatchDev b, uint32_t k)
{
    const ViewDev& v = batch_view(b, blockIdx.y);
    if (k < v.hiz_tile_levels || k >= v.hiz_tail_start) return;
    const PyramidDesc& pd = v.pyr;
    const uint32_t sw = pd.w[k - 1], sh = pd.h[k - 1], dw = pd.w[k], dh = pd.h[k];
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= dw * dh) return;
    const float* src = v.pyramid + pd.off[k - 1];
    const uint32_t x = t % dw, y = t / dw;
    int x0, x1, y0, y1;
    footprint(x, dw, sw, x0, x1);
    footprint(y, dh, sh, y0, y1);
    float d = src[y0 * sw + x0];
    d = gmin(d, src[y0 * sw + x1]);
    d = gmin(d, src[y1 * sw + x0]);
    d = gmin(d, src[y1 * sw + x1]);
    v.pyramid[pd.off[k] + t] = d;
}

} // namespace

// First level of the tail: the first level after the tile pass small enough
// for the tail's LDS buffer.
uint32_t hiz_tail_start(const PyramidDesc& pd, uint32_t tile_levels)
{
    uint32_t s = tile_levels;
    while (s < pd.levels && (uint64_t)pd.w[s] * pd.h[s] > kHizTailTexels) ++s;
    return s;
}

// Number of levels the tile kernel can produce: level 0 plus every following
// level whose source is exactly twice its size, at most 6 (32 -> 1).
uint32_t hiz_tile_levels(const PyramidDesc& pd)
{
    uint32_t n = 1;
    while (n < pd.levels && n < 6 && pd.w[n - 1] == 2 * pd.w[n] && pd.h[n - 1] == 2 * pd.h[n]) ++n;
    return n;
}

void launch_hiz(const BatchDev& b, const BatchGrid& g, hipStream_t st)
{
    if (g.hiz_blocks) hiz_tiles_kernel<<<g.hiz_blocks, 256, 0, st>>>(b);
    for (uint32_t k = 1; k < kHizMaxLevels; ++k)
        if (g.hiz_level_blocks[k]) hiz_level_kernel<<<dim3(g.hiz_level_blocks[k], g.nviews), 256, 0, st>>>(b, k);
    if (g.max_tail_levels) hiz_tail_kernel<<<g.nviews, 1024, 0, st>>>(b);
}

} // namespace vkgv
