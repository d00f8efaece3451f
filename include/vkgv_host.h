/*
 * vkgv_host.h - C ABI of libvkgv_host.so: the host-side producers of the
 * frame's inputs, re-implemented in C++ (no GPU code):
 *   - camera math (Camera::updateCamera, camera.cpp:108-194),
 *   - the scene/draw interface (World::rebuildDrawBuffer and
 *     World::updateTransformBuffer, world.cpp:187-345),
 *   - a meshlet builder at the reference's limits (64 vertices, 124
 *     triangles, assets.cpp:322-373), with per-meshlet normal cones,
 *   - seeded synthetic scenes for the benchmark configurations,
 *   - EXT_meshopt_compression decoders for glTF buffer views.
 * Arrays returned by accessors stay valid until vkgv_scene_free().
 */
#ifndef VKGV_HOST_H
#define VKGV_HOST_H

#include "vkgv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A camera as the application sets it up (camera.cpp:170-193). */
typedef struct vkgv_view_desc {
    uint32_t width, height;
    uint32_t orthographic;           /* 0: perspectiveRH_ZO, 1: orthoRH_ZO */
    uint32_t pad;
    float    eye[3];
    float    target[3];
    float    up[3];
    float    fovy;                   /* radians (perspective) */
    float    znear, zfar;
    float    ortho_half_width, ortho_half_height;
} vkgv_view_desc;

/* Builds the camera buffer of a first frame: P[1][1] *= -1,
 * VP = reverseDepth(P) * lookAt, frustum planes, prev* = current. */
void vkgv_camera_from_view(const vkgv_view_desc* view, vkgv_camera* out);
/* Next frame: prev* <- current, then the new matrices (camera.cpp:181-193). */
void vkgv_camera_advance(const vkgv_camera* prev, const vkgv_view_desc* view, vkgv_camera* out);

/* ---- meshlets ---------------------------------------------------------- */
typedef struct vkgv_meshlet_build {
    uint32_t*          vertex_indices;   uint64_t vertex_index_count;
    uint8_t*           triangle_indices; uint64_t triangle_index_bytes;
    vkgv_meshlet*      meshlets;         uint32_t meshlet_count;
    vkgv_meshlet_cone* cones;
} vkgv_meshlet_build;

/* Partitions an indexed triangle list into meshlets of at most max_vertices
 * vertices and max_triangles triangles (max_triangles a multiple of 4, at
 * most 124), fills AABBs as assets.cpp:349-372 and normal cones.  Returns 0
 * on success; free the result with vkgv_meshlet_build_free. */
int  vkgv_build_meshlets(const float* positions, uint32_t vertex_count,
                         const uint32_t* indices, uint64_t index_count,
                         uint32_t max_vertices, uint32_t max_triangles,
                         vkgv_meshlet_build* out);
void vkgv_meshlet_build_free(vkgv_meshlet_build* b);

/* ---- scenes ------------------------------------------------------------ */
typedef struct vkgv_scene vkgv_scene;

/* Scene configurations:
 *  1  256 icospheres (subdivision 4) at random in a 100 m cube, 1280x720
 *  2  occlusion-heavy city: 64x64 boxes + noise-displaced int16-quantized
 *     spheres (variant 0: 2048 subdivision-4 spheres on a checkerboard,
 *     about 10.5M triangles; variant 1: 4096 subdivision-5, 83.9M), 3840x2160
 *  3  configuration 1's scene seen by 64 views (4 ortho cascades 2048^2,
 *     6 cube faces 512^2, 54 random perspective views 256..1920 x 256..1080)
 *  4  one ~50k-vertex icosphere-derived mesh with per-vertex albedo, 8 ring
 *     cameras at 512x512
 *  5  uneven stress: 200 meshes with log-uniform meshlet counts (variant
 *     scales the count down: meshes = 200 >> variant), mixed quantization,
 *     1-64 instances, 16 views up to 3840x2160 */
vkgv_scene* vkgv_scene_generate(uint32_t config, uint64_t seed, uint32_t variant);
void        vkgv_scene_free(vkgv_scene* s);
const char* vkgv_scene_error(void);

uint32_t                   vkgv_scene_mesh_count(const vkgv_scene* s);
const vkgv_mesh_desc*      vkgv_scene_meshes(const vkgv_scene* s);
uint32_t                   vkgv_scene_primitive_count(const vkgv_scene* s);
const vkgv_primitive_desc* vkgv_scene_primitives(const vkgv_scene* s);
uint32_t                   vkgv_scene_material_count(const vkgv_scene* s);
const vkgv_material*       vkgv_scene_materials(const vkgv_scene* s);
uint32_t                   vkgv_scene_draw_count(const vkgv_scene* s);
const vkgv_meshlet_draw*   vkgv_scene_draws(const vkgv_scene* s);
uint32_t                   vkgv_scene_transform_count(const vkgv_scene* s);
const float*               vkgv_scene_transforms(const vkgv_scene* s);
uint32_t                   vkgv_scene_view_count(const vkgv_scene* s);
const vkgv_view_desc*      vkgv_scene_views(const vkgv_scene* s);
uint64_t                   vkgv_scene_triangle_count(const vkgv_scene* s);  /* sum over draws */

/* ---- World: node hierarchy -> draws + transforms ---------------------- */
/* (vkgv_node is declared in vkgv.h, which the device-side transform update shares.) */

/* Walks the roots depth-first (World::iterateNode, world.cpp:187-228) with
 * M = parent * T * R * S; every node with a mesh takes the next transform
 * index and emits one draw per (primitive of the mesh, meshlet)
 * (world.cpp:230-293); transforms are written in the same order
 * (world.cpp:295-345).  mesh_primitives[m] lists the primitives of glTF mesh
 * m as (first, count) into prim_list; meshlet_counts[p] is the meshlet count
 * of primitive p.  Pass NULL outputs to query the counts. */
int vkgv_world_build(const vkgv_node* nodes, uint32_t node_count,
                     const uint32_t* children, const uint32_t* roots, uint32_t root_count,
                     const uint32_t* mesh_primitives, const uint32_t* prim_list,
                     const uint32_t* meshlet_counts,
                     vkgv_meshlet_draw* draws, uint32_t* draw_count,
                     float* transforms, uint32_t* transform_count);

/* ---- EXT_meshopt_compression ------------------------------------------- */

/* Decoders of compressed glTF buffer views, replacing the meshoptimizer calls
 * of CompressedBufferDataAdapter::ExecuteRange (assets.cpp:111-171):
 * meshopt_decodeVertexBuffer (assets.cpp:132-133), meshopt_decodeIndexBuffer
 * (:137-138), meshopt_decodeIndexSequence (:142-143) and
 * meshopt_decodeFilterOct / Quat / Exp (:152-167), with the same arguments
 * and return codes: 0 success, -1 bad header, version or argument, -2 the
 * stream ends early, -3 trailing data.  Vertex streams: codec version 0
 * (0xA0), vertex_size a multiple of 4 up to 256.  Triangle streams: versions 0
 * and 1 (0xE0, 0xE1), index_size 2 or 4.  Index sequences: 0xD0 / 0xD1. */
int vkgv_meshopt_decode_vertex_buffer(void* destination, size_t vertex_count, size_t vertex_size,
                                      const unsigned char* buffer, size_t buffer_size);
int vkgv_meshopt_decode_index_buffer(void* destination, size_t index_count, size_t index_size,
                                     const unsigned char* buffer, size_t buffer_size);
int vkgv_meshopt_decode_index_sequence(void* destination, size_t index_count, size_t index_size,
                                       const unsigned char* buffer, size_t buffer_size);
/* In-place filters over count elements of `stride` bytes: octahedral
 * normals (stride 4: int8, 8: int16), quaternions (stride 8), exponential
 * floats (stride a multiple of 4). */
void vkgv_meshopt_decode_filter_oct(void* buffer, size_t count, size_t stride);
void vkgv_meshopt_decode_filter_quat(void* buffer, size_t count, size_t stride);
void vkgv_meshopt_decode_filter_exp(void* buffer, size_t count, size_t stride);

enum { VKGV_MESHOPT_ATTRIBUTES = 0, VKGV_MESHOPT_TRIANGLES = 1, VKGV_MESHOPT_INDICES = 2 };
enum { VKGV_MESHOPT_FILTER_NONE = 0, VKGV_MESHOPT_FILTER_OCTAHEDRAL = 1, VKGV_MESHOPT_FILTER_QUATERNION = 2,
       VKGV_MESHOPT_FILTER_EXPONENTIAL = 3 };
/* One buffer view as the adapter decodes it (assets.cpp:111-171): the mode's
 * decoder into count * stride bytes, then the filter (ATTRIBUTES only; the
 * stride must suit the filter).  Returns the decoder's code, or -1 for an
 * invalid mode / filter combination. */
int vkgv_meshopt_decode(void* destination, size_t count, size_t stride, const unsigned char* buffer,
                        size_t buffer_size, int mode, int filter);

#ifdef __cplusplus
}
#endif
#endif
