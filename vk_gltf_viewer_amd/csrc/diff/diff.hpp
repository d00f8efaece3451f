// Launchers of the differentiable form's kernels (diff.hip).  Plain pointers;
// the torch bindings (diff_ext.cpp) own the tensors and the autograd graph.
#pragma once
#include <hip/hip_runtime.h>

namespace vkgv_diff {

// keys: C*H*W u64 set to all ones by the caller; ids / depth: C*H*W.
void raster_tris(const float* clip, const int* tris, int C, int I, int V, int T, int H, int W, int cull_back,
                 unsigned long long* keys, int* ids, float* depth, hipStream_t st);
// Interpolation and antialiasing in float or double (clip, attributes,
// images and gradients share the type; depth stays float).
#define VKGV_DIFF_DECLARE(R)                                                                                          \
    void interp_fwd(const R* clip, const R* attr, const int* tris, const int* ids, int C, int I, int V, int T, int H,  \
                    int W, int A, R* out, hipStream_t st);                                                             \
    /* g_clip and g_attr are accumulated into (zeroed by the caller) */                                               \
    void interp_bwd(const R* clip, const R* attr, const int* tris, const int* ids, int C, int I, int V, int T, int H,  \
                    int W, int A, const R* g_out, R* g_clip, R* g_attr, hipStream_t st);                               \
    void aa_fwd(const R* color, const R* clip, const int* tris, const int* ids, const float* depth, int C, int I,      \
                int V, int T, int H, int W, int A, R* out, hipStream_t st);                                            \
    /* g_color is written; g_clip is accumulated into */                                                              \
    void aa_bwd(const R* color, const R* clip, const int* tris, const int* ids, const float* depth, int C, int I,      \
                int V, int T, int H, int W, int A, const R* g_out, R* g_color, R* g_clip, hipStream_t st);

VKGV_DIFF_DECLARE(float)
VKGV_DIFF_DECLARE(double)
#undef VKGV_DIFF_DECLARE

} // namespace vkgv_diff
