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

// Interpolation and antialiasing.  C is the type of the clip positions and
// their gradient, R that of attributes, images, their gradients and the
// arithmetic: (float, float), (double, double) or (double, float).  Depth
// stays float.
#define VKGV_DIFF_DECLARE(C, R)                                                                                       \
    void interp_fwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, R* out, hipStream_t st);                                                             \
    /* g_clip and g_attr are accumulated into (zeroed by the caller) */                                               \
    void interp_bwd(const C* clip, const R* attr, const int* tris, const int* ids, int nc, int I, int V, int T, int H, \
                    int W, int A, const R* g_out, C* g_clip, R* g_attr, hipStream_t st);                               \
    void aa_fwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, R* out, hipStream_t st);                                            \
    /* g_color is written; g_clip is accumulated into */                                                              \
    void aa_bwd(const R* color, const C* clip, const int* tris, const int* ids, const float* depth, int nc, int I,     \
                int V, int T, int H, int W, int A, const R* g_out, R* g_color, C* g_clip, hipStream_t st);

VKGV_DIFF_DECLARE(float, float)
VKGV_DIFF_DECLARE(double, double)
VKGV_DIFF_DECLARE(double, float)
#undef VKGV_DIFF_DECLARE

} // namespace vkgv_diff
