/*
 * cpuref.h - the oracle: a scalar C restatement of vk_gltf_viewer's GPU-driven
 * frame, used ONLY as test infrastructure (tests/, the smoke check and the
 * bench's CPU-baseline leg).  Nothing in the product links or calls it.
 *
 * Each function cites the reference file:line whose semantics it restates.
 * Floating-point evaluation order is fixed and contraction is off
 * (-ffp-contract=off), so that a device implementation using the same order
 * can be compared bit for bit.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   camera math        closed forms (glm perspectiveRH_ZO / lookAtRH), frustum
 *                      corners of the inverse VP lie on the planes
 *   frustum / AABB     brute force on corner sets, Arvo bound invariant
 *   projectAabb/probe  hand-computed pyramids (all-far, all-near, single texel)
 *   Hi-Z reduce        brute-force 2x2 min on even sizes, footprint formula on odd
 *   raster             brute-force float edge tests on tiny scenes, watertight
 *                      shared edges, coverage area vs. analytic polygon area
 *   resolve            IEC 61966-2-1 sRGB encode
 * The reference itself (Vulkan, lavapipe, glslang, submodules) cannot be built
 * in this environment, and it ships no tests or golden vectors.
 */
#ifndef VKGV_CPUREF_H
#define VKGV_CPUREF_H

#include <stdint.h>
#include "../include/vkgv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A scene as the oracle sees it: host arrays in the contract layout. */
typedef struct ref_scene {
    const vkgv_mesh_desc*      meshes;
    uint32_t                   mesh_count;
    const vkgv_primitive_desc* primitives;
    uint32_t                   primitive_count;
    const vkgv_material*       materials;
    uint32_t                   material_count;
    const vkgv_meshlet_draw*   draws;
    uint32_t                   draw_count;
    const float*               transforms;      /* mat4, column-major */
    uint32_t                   transform_count;
    const vkgv_texture_desc*   textures;        /* albedo textures (VKGV_RESOLVE_TEXTURED) */
    uint32_t                   texture_count;
} ref_scene;

/* ---- camera.cpp --------------------------------------------------------- */
void ref_mat4_mul(const float* a, const float* b, float* out);
void ref_mat4_vec4(const float* m, const float* v, float* out);   /* m * v */
void ref_cross3(const float* a, const float* b, float* out);
void ref_perspective_rh_zo(float fovy, float aspect, float znear, float zfar, float* out);
void ref_ortho_rh_zo(float l, float r, float b, float t, float znear, float zfar, float* out);
void ref_look_at_rh(const float* eye, const float* center, const float* up, float* out);
void ref_reverse_depth(const float* proj, float* out);
void ref_frustum_planes(const float* vp, float* planes /* 6*4 */);
void ref_make_camera(const float* view, const float* proj_no_flip, vkgv_camera* out);

/* ---- culling.h.glsl / visbuffer.task.glsl ------------------------------- */
int  ref_aabb_in_frustum(const float* center, const float* extents, const float* planes);
void ref_world_aabb(const float* m, const float* c, const float* e, float* wc, float* we);
void ref_project_aabb(const float* c, const float* e, const float* vp, float* ss_min, float* ss_max);

/* Hi-Z pyramid geometry (application.cpp:472-529). Returns the level count. */
uint32_t ref_pyramid_levels(uint32_t w, uint32_t h, uint32_t* sizes, uint64_t* offsets);
/* textureLod with the MIN-reduction linear sampler (application.cpp:437-458). */
float ref_pyramid_probe(const float* pyramid, uint32_t w, uint32_t h, const float* uv, float level);
/* floor(log2(x)) as the probe uses it. */
float ref_floor_log2(float x);
/* Occlusion test of one world AABB against a pyramid (visbuffer.task.glsl:54-65). */
int ref_occlusion_visible(const float* wc, const float* we, const float* vp,
                          const float* pyramid, uint32_t w, uint32_t h);
/* Normal-cone back-face test (extension).  Returns 1 if the cone culls. */
int ref_cone_culled(const vkgv_meshlet_cone* cone, const float* m, const float* eye);

/* Per-draw visibility of the single-pass reference cull: frustum, then
 * occlusion against `pyramid` with prevOcclusionViewProjection.
 * visible[d] = 0/1.  flags: VKGV_CULL_* bits. */
void ref_cull(const ref_scene* s, const vkgv_camera* cam, const float* eye,
              const float* pyramid, uint32_t w, uint32_t h, uint32_t flags,
              uint8_t* visible);

/* ---- hiz_reduce.comp.glsl ----------------------------------------------- */
void ref_hiz_build(const float* depth, uint32_t w, uint32_t h, float* pyramid);

/* ---- visbuffer.mesh.glsl + fixed-function raster + visbuffer.frag.glsl --- */
/* Raster statistics accumulated by ref_raster (all threads) since the last
 * reset: triangles past the back-face test, and fragments (covered pixel
 * centres with 0 <= z <= 1).  reset != 0 zeroes them after reading. */
void ref_raster_stats(int reset, uint64_t* tris, uint64_t* frags);
/* OpenMP threads of the frame functions (n <= 0: unchanged); returns the count. */
int  ref_threads(int n);
void ref_keys_clear(uint64_t* keys, uint64_t n);
/* Rasterize the listed draws (ascending order is not required: the key
 * minimum is order independent). */
void ref_raster(const ref_scene* s, const float* vp, const uint32_t* draw_list, uint32_t n,
                uint32_t w, uint32_t h, uint32_t flags, uint64_t* keys);
/* Rasterize one clip-space triangle (x,y,z,w per vertex) with id. */
void ref_raster_triangle(const float* c0, const float* c1, const float* c2, uint32_t id,
                         uint32_t w, uint32_t h, uint64_t* keys);
void ref_keys_depth(const uint64_t* keys, uint64_t n, float* depth);
void ref_keys_ids(const uint64_t* keys, uint64_t n, uint32_t* ids);

/* ---- visbuffer_resolve.comp.glsl + srgb.h.glsl -------------------------- */
float ref_from_linear(float c);
void  ref_srgb_codes(const float* c, uint64_t n, uint8_t* out);
void  ref_resolve(const ref_scene* s, const float* vp, const uint64_t* keys,
                  uint32_t w, uint32_t h, uint32_t flags, uint32_t* rgba);

/* ---- whole frames -------------------------------------------------------- */
/* Reference frame: cull vs. prev pyramid, raster, resolve, rebuild pyramid.
 * pyramid is read (previous frame) and then overwritten (this frame). */
uint32_t ref_frame_reference(const ref_scene* s, const vkgv_camera* cam, const float* eye,
                             uint32_t w, uint32_t h, uint32_t flags,
                             float* pyramid, uint64_t* keys, uint32_t* rgba,
                             uint8_t* visible);
/* Two-pass frame.  history[d] (0/1) is read and updated; visible_early and
 * visible_late receive per-draw flags of the two raster passes; pyramid
 * receives the mid-frame pyramid. */
void ref_frame_two_pass(const ref_scene* s, const vkgv_camera* cam, const float* eye,
                        uint32_t w, uint32_t h, uint32_t flags,
                        uint8_t* history, float* pyramid, uint64_t* keys, uint32_t* rgba,
                        uint8_t* visible_early, uint8_t* visible_late);

#ifdef __cplusplus
}
#endif
/* ---- anim.c: transform update and motion vectors ------------------------ */
double ref_series_acos(double x);
double ref_series_sin(double a);
void ref_anim_sample(const vkgv_anim_sampler* sp, const float* times, const float* values, int n, float time,
                     float* out);
int  ref_update_transforms(const vkgv_node* nodes, uint32_t node_count, const uint32_t* children,
                           const uint32_t* roots, uint32_t root_count, const vkgv_anim_sampler* samplers,
                           const float* times, const float* values, const vkgv_anim_channel* channels,
                           uint32_t channel_count, float time, float* out, uint32_t* count);
void ref_motion_vectors(const ref_scene* s, const vkgv_camera* cam, const float* prev_transforms,
                        const uint64_t* keys, uint32_t w, uint32_t h, float* motion);

#endif
