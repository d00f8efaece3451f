// Launchers of the differentiable form's kernels (diff.hip).  Plain pointers;
// the torch bindings (diff_ext.cpp) own the tensors and the autograd graph.
#pragma once
#include <hip/hip_runtime.h>

namespace vkgv_diff {

// keys: C*H*W u64 set to all ones by the caller; ids / depth: C*H*W.
void raster_tris(const float* clip, const int* tris, int C, int I, int V, int T, int H, int W, int cull_back,
                 unsigned long long* keys, int* ids, float* depth, hipStream_t st);
// Vertex transform: clip [CI, V, 4] = A [CI, 4, 4] (pos [V, 3], 1), and its
// backward: g_pos [V, 3] (written) and g_A [CI, 4, 4] (accumulated into,
// zeroed by the caller).
void transform_fwd(const float* pos, const float* A, int CI, int V, float* clip, hipStream_t st);
void transform_fwd(const double* pos, const double* A, int CI, int V, double* clip, hipStream_t st);
void transform_bwd(const float* g_clip, const float* pos, const float* A, int CI, int V, float* g_pos, float* g_A,
                   hipStream_t st);
void transform_bwd(const double* g_clip, const double* pos, const double* A, int CI, int V, double* g_pos,
                   double* g_A, hipStream_t st);

// Vertex adjacency of the mesh tris [T, 3] over V vertices, for the
// backward gathers: off [2 V + 1] (off[0..V] the CSR offsets, the rest
// scratch), adj [3 T] (corner indices 3 t + k, by vertex, increasing).
void mesh_adjacency(const int* tris, int T, int V, int* off, int* adj, hipStream_t st);

// Screen boxes of every (camera, instance, triangle), for the stray flags of
// the forwards and the gathers of the backwards: tri_boxes_words(nc, I, T)
// u64 entries (nc * I * T box words, then one flag byte per triangle that the
// forwards set for the triangles with a gathered pixel).
size_t tri_boxes_words(int nc, int I, int T);
void tri_boxes(const float* clip, const int* tris, int nc, int I, int V, int T, int H, int W,
               unsigned long long* boxes, hipStream_t st);
void tri_boxes(const double* clip, const int* tris, int nc, int I, int V, int T, int H, int W,
               unsigned long long* boxes, hipStream_t st);

// Entries of the per-block stray flags that the forwards write and the
// backwards read (one per 256 pixels).
size_t stray_flag_count(int nc, int H, int W);

// Interpolation and antialiasing.  C is the type of the clip positions and
// their gradient, R that of attributes, images, their gradients and the
// arithmetic: (float, float), (double, double) or (double, float).  Depth
// stays float.  The forwards take the boxes of the same clip positions and
// write stray_blocks (aa_fwd: unless null); the backwards take them with the
// adjacency of tris and write every output (g_clip with z = 0).
// shade_bwd is the backward of aa(interp(...)) in one: g_color is scratch.
// Scratch: contrib (contrib_i), nc * H * W * 3 * (3 + A) entries for the
// interpolation and contrib (contrib_a), nc * H * W * 9 for the
// antialiasing; part, nc * I * V * A entries.
#define VKGV_DIFF_DECLARE(C, R)                                                                                       \
    void interp_fwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks, hipStream_t st);          \
    void interp_bwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks, const int* adj_off,   \
                    const int* adj, const R* g_out, R* contrib, R* part, C* g_clip, R* g_attr, hipStream_t st);        \
    void aa_fwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, R* out, unsigned long long* boxes, unsigned* stray_blocks,          \
                hipStream_t st);                                                                                       \
    void aa_bwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, const unsigned long long* boxes, const unsigned* stray_blocks,      \
                const int* adj_off, const int* adj, const R* g_out, R* contrib, R* g_color, C* g_clip,                 \
                hipStream_t st);                                                                                       \
    void shade_bwd(const R* color, const C* clip, const R* attr, const int* tris, const int* ids, const float* depth,  \
                   int nc, int I, int V, int T, int H, int W, int A, const unsigned long long* boxes,                  \
                   const unsigned* stray_blocks, const int* adj_off, const int* adj, const R* g_out, R* g_color,       \
                   R* contrib_a, R* contrib_i, R* part, C* g_clip, R* g_attr, hipStream_t st);

VKGV_DIFF_DECLARE(float, float)
VKGV_DIFF_DECLARE(double, double)
VKGV_DIFF_DECLARE(double, float)
#undef VKGV_DIFF_DECLARE

} // namespace vkgv_diff
