/*
 * vkgv.h - C ABI of libvkgv_hip.so, the MI355X (gfx950) implementation of the
 * GPU-driven frame of vk_gltf_viewer: meshlet cull + compaction, Hi-Z depth
 * pyramid, visibility-buffer rasterization and material resolve.
 *
 * Everything crossing this boundary is plain C: fixed-layout structs, pointers
 * and sizes. No C++ types, no exceptions, no HIP types (a stream is a void*).
 *
 * The data contract mirrors the reference's dual-compiled GLSL/C++ structs
 * byte for byte (shaders/mesh_common.h.glsl, shaders/visbuffer/visbuffer.h.glsl),
 * so that a caller holding the reference's glsl:: arrays can hand them over
 * unchanged.  Sizes and offsets are pinned by static assertions below.
 *
 * Status codes: 0 = ok, negative = error; vkgv_last_error() returns the text of
 * the last error raised on a context.  A context and its views are externally
 * synchronized: one thread per context per call.
 */
#ifndef VKGV_H
#define VKGV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Limits (shaders/mesh_common.h.glsl:36-38, shaders/visbuffer/visbuffer.h.glsl:15-17) */
#define VKGV_MAX_VERTICES            64u   /* maxVertices */
#define VKGV_MAX_PRIMITIVES          126u  /* maxPrimitives */
#define VKGV_MESHLET_TRIANGLES       124u  /* alignDown(maxPrimitives, 4): assets.cpp:324 */
#define VKGV_MAX_MESHLETS_PER_TASK   102u  /* maxMeshlets */
#define VKGV_TRIANGLE_BITS           7u    /* triangleBits */
#define VKGV_DRAW_INDEX_BITS         25u   /* drawIndexBits */
#define VKGV_MAX_DRAWS               (1u << VKGV_DRAW_INDEX_BITS)
#define VKGV_VISBUFFER_CLEAR         0xFFFFFFFFu /* visbufferClearValue */

/* ------------------------------------------------------------------------ */
/* Data contract.  Column-major float[16] matrices (glm::mat4 layout).       */

/* glsl::Meshlet, mesh_common.h.glsl:49-58 (36 bytes) */
typedef struct vkgv_meshlet {
    uint32_t vertex_offset;    /* into the primitive's vertex-index buffer */
    uint32_t triangle_offset;  /* byte offset into the primitive's u8 triangle buffer */
    uint8_t  vertex_count;     /* <= 64 */
    uint8_t  triangle_count;   /* <= 124 */
    uint8_t  pad[2];
    float    aabb_extents[3];  /* @12 */
    float    aabb_center[3];   /* @24 */
} vkgv_meshlet;

/* glsl::Vertex, mesh_common.h.glsl:60-71 (24 bytes) */
typedef struct vkgv_vertex {
    float    position[3];      /* @0 */
    uint8_t  color[4];         /* @12, unorm8 RGBA */
    uint8_t  normal[3];        /* @16, (n + 1) * 127 */
    uint8_t  pad;
    uint16_t uv[2];            /* @20, float16 bits */
} vkgv_vertex;

/* glsl::MeshletDraw, mesh_common.h.glsl:96-100 (12 bytes) */
typedef struct vkgv_meshlet_draw {
    uint32_t primitive_index;
    uint32_t meshlet_index;
    uint32_t transform_index;
} vkgv_meshlet_draw;

/* glsl::Primitive, mesh_common.h.glsl:102-113 (64 bytes).  The four buffer
 * fields are device addresses, exactly as the reference's buffer-device
 * addresses; the library fills them in vkgv_upload_scene. */
typedef struct vkgv_primitive {
    uint64_t vertex_index_buffer;    /* uint32_t[]            */
    uint64_t primitive_index_buffer; /* uint8_t[] (3 per tri) */
    uint64_t vertex_buffer;          /* vkgv_vertex[]         */
    uint64_t meshlet_buffer;         /* vkgv_meshlet[]        */
    float    aabb_extents[3];        /* @32 */
    float    aabb_center[3];         /* @44 */
    uint32_t meshlet_count;          /* @56 */
    uint32_t material_index;         /* @60, glTF index + 1; 0 = default material */
} vkgv_primitive;

/* glsl::Material, mesh_common.h.glsl:115-126 (48 bytes) */
typedef struct vkgv_material {
    float    albedo_factor[4];       /* linear RGBA */
    uint32_t albedo_index;           /* @16, texture handle: index into vkgv_upload_textures, ~0u = none */
    float    uv_offset[2];           /* @20 */
    float    uv_scale[2];            /* @28 */
    float    uv_rotation;            /* @36 */
    float    alpha_cutoff;           /* @40 */
    uint32_t double_sided;           /* @44, GLSL bool (4 bytes) */
} vkgv_material;

/* glsl::Camera, mesh_common.h.glsl:20-27 (352 bytes) */
typedef struct vkgv_camera {
    float prev_view_projection[16];
    float prev_occlusion_view_projection[16];
    float view_projection[16];
    float occlusion_view_projection[16];
    float frustum[6][4];             /* normalized planes, w negated (camera.cpp:70-84) */
} vkgv_camera;

/* Normal cone of a meshlet (extension: the reference leaves cone culling off,
 * assets.cpp:323).  Layout of meshopt_Bounds' cone part: the meshlet is
 * back-facing for every eye position p with
 *     dot(normalize(apex - p), axis) >= cutoff      (cutoff >= 1: never). */
typedef struct vkgv_meshlet_cone {
    float apex[3];
    float axis[3];
    float cutoff;
    float pad;
} vkgv_meshlet_cone;                 /* 32 bytes */

/* The contract layouts (sizes and offsets of the reference's scalar-layout
 * GLSL structs, mesh_common.h.glsl:20-126), checked by every compiler that
 * includes this header. */
#ifdef __cplusplus
#define VKGV_LAYOUT(cond, what) static_assert(cond, what)
#else
#define VKGV_LAYOUT(cond, what) _Static_assert(cond, what)
#endif
VKGV_LAYOUT(sizeof(vkgv_meshlet) == 36, "Meshlet is 36 bytes");
VKGV_LAYOUT(offsetof(vkgv_meshlet, triangle_offset) == 4, "Meshlet.triangleOffset @4");
VKGV_LAYOUT(offsetof(vkgv_meshlet, vertex_count) == 8, "Meshlet.vertexCount @8");
VKGV_LAYOUT(offsetof(vkgv_meshlet, triangle_count) == 9, "Meshlet.triangleCount @9");
VKGV_LAYOUT(offsetof(vkgv_meshlet, aabb_extents) == 12, "Meshlet.aabbExtents @12");
VKGV_LAYOUT(offsetof(vkgv_meshlet, aabb_center) == 24, "Meshlet.aabbCenter @24");
VKGV_LAYOUT(sizeof(vkgv_vertex) == 24, "Vertex is 24 bytes");
VKGV_LAYOUT(offsetof(vkgv_vertex, color) == 12, "Vertex.color @12");
VKGV_LAYOUT(offsetof(vkgv_vertex, normal) == 16, "Vertex.normal @16");
VKGV_LAYOUT(offsetof(vkgv_vertex, uv) == 20, "Vertex.uv @20");
VKGV_LAYOUT(sizeof(vkgv_meshlet_draw) == 12, "MeshletDraw is 12 bytes");
VKGV_LAYOUT(offsetof(vkgv_meshlet_draw, meshlet_index) == 4, "MeshletDraw.meshletIndex @4");
VKGV_LAYOUT(offsetof(vkgv_meshlet_draw, transform_index) == 8, "MeshletDraw.transformIndex @8");
VKGV_LAYOUT(sizeof(vkgv_primitive) == 64, "Primitive is 64 bytes");
VKGV_LAYOUT(offsetof(vkgv_primitive, meshlet_buffer) == 24, "Primitive.meshletBuffer @24");
VKGV_LAYOUT(offsetof(vkgv_primitive, aabb_extents) == 32, "Primitive.aabbExtents @32");
VKGV_LAYOUT(offsetof(vkgv_primitive, aabb_center) == 44, "Primitive.aabbCenter @44");
VKGV_LAYOUT(offsetof(vkgv_primitive, meshlet_count) == 56, "Primitive.meshletCount @56");
VKGV_LAYOUT(offsetof(vkgv_primitive, material_index) == 60, "Primitive.materialIndex @60");
VKGV_LAYOUT(sizeof(vkgv_material) == 48, "Material is 48 bytes");
VKGV_LAYOUT(offsetof(vkgv_material, albedo_index) == 16, "Material.albedoIndex @16");
VKGV_LAYOUT(offsetof(vkgv_material, uv_offset) == 20, "Material.uvOffset @20");
VKGV_LAYOUT(offsetof(vkgv_material, uv_scale) == 28, "Material.uvScale @28");
VKGV_LAYOUT(offsetof(vkgv_material, uv_rotation) == 36, "Material.uvRotation @36");
VKGV_LAYOUT(offsetof(vkgv_material, alpha_cutoff) == 40, "Material.alphaCutoff @40");
VKGV_LAYOUT(offsetof(vkgv_material, double_sided) == 44, "Material.doubleSided @44");
VKGV_LAYOUT(sizeof(vkgv_camera) == 352, "Camera is 352 bytes");
VKGV_LAYOUT(offsetof(vkgv_camera, prev_occlusion_view_projection) == 64, "Camera.prevOcclusionViewProjection @64");
VKGV_LAYOUT(offsetof(vkgv_camera, view_projection) == 128, "Camera.viewProjection @128");
VKGV_LAYOUT(offsetof(vkgv_camera, occlusion_view_projection) == 192, "Camera.occlusionViewProjection @192");
VKGV_LAYOUT(offsetof(vkgv_camera, frustum) == 256, "Camera.frustum @256");
VKGV_LAYOUT(sizeof(vkgv_meshlet_cone) == 32, "meshlet cone is 32 bytes");

/* ------------------------------------------------------------------------ */
/* Scene upload (assets.cpp:375-429 + world.cpp:10-185 equivalent).          */

/* One set of per-primitive buffers (PrimitiveBuffers, assets.hpp:16-23).  A
 * mesh may be shared by several primitives (e.g. one icosphere drawn with 256
 * materials); each primitive's four device addresses then point at it. */
typedef struct vkgv_mesh_desc {
    const uint32_t*          vertex_indices;     /* meshlet-local -> vertex */
    uint64_t                 vertex_index_count;
    const uint8_t*           triangle_indices;   /* 3 u8 per triangle, 4-byte padded per meshlet */
    uint64_t                 triangle_index_bytes;
    const vkgv_vertex*       vertices;
    uint64_t                 vertex_count;
    const vkgv_meshlet*      meshlets;
    uint32_t                 meshlet_count;
    uint32_t                 pad;
    const vkgv_meshlet_cone* cones;              /* optional, meshlet_count entries */
} vkgv_mesh_desc;

typedef struct vkgv_primitive_desc {
    uint32_t mesh_index;                 /* into the vkgv_mesh_desc array */
    uint32_t material_index;             /* glTF index + 1; 0 = default */
    float    aabb_extents[3];
    float    aabb_center[3];
} vkgv_primitive_desc;

/* ------------------------------------------------------------------------ */
/* Frame flags                                                               */
#define VKGV_CULL_FRUSTUM        (1u << 0)  /* isAabbInFrustum (culling.h.glsl:8-19) */
#define VKGV_CULL_OCCLUSION_PREV (1u << 1)  /* single pass vs. the previous frame's pyramid (visbuffer.task.glsl:54-65) */
#define VKGV_CULL_TWO_PASS       (1u << 2)  /* early pass = visible last frame, pyramid, late pass vs. it */
#define VKGV_CULL_CONE           (1u << 3)  /* normal-cone back-face cull per meshlet */
#define VKGV_RESOLVE_BARY        (1u << 4)  /* resolve recomputes barycentrics and interpolates vertex colour */
#define VKGV_BUILD_PYRAMID       (1u << 5)  /* build the pyramid after the resolve, for the next frame */
#define VKGV_TIMING              (1u << 6)  /* record per-pass HIP events (vkgv_view_read_stats) */
#define VKGV_NO_BACKFACE         (1u << 7)  /* disable per-triangle back-face culling */
#define VKGV_SKIP_RESOLVE        (1u << 8)  /* leave the colour image untouched */
#define VKGV_GRAPH               (1u << 9)  /* vkgv_render_batch: replay the batch's frame as one captured hipGraph */
#define VKGV_RESOLVE_TEXTURED    (1u << 10) /* as RESOLVE_BARY, times the material's albedo texture sampled at transformUv(uv) */
#define VKGV_MOTION_VECTORS      (1u << 11) /* resolve also writes per-pixel motion vectors (visbuffer.frag.glsl:38) */
#define VKGV_STATS               (1u << 12) /* count rasterized triangles, fragments and visibility-buffer atomics
                                               (vkgv_frame_stats); counting kernels are separate instantiations,
                                               frames without the flag run uncounted code */

/* The reference's own frame: frustum + previous-frame Hi-Z occlusion, pyramid
 * rebuilt after the resolve (application.cpp:843-1003). */
#define VKGV_MODE_REFERENCE (VKGV_CULL_FRUSTUM | VKGV_CULL_OCCLUSION_PREV | VKGV_BUILD_PYRAMID)

typedef struct vkgv_frame_desc {
    vkgv_camera camera;              /* the per-frame camera buffer (camera.cpp:180-193) */
    float       eye[3];              /* world-space eye position: cone culling only */
    uint32_t    flags;
    void*       stream;              /* hipStream_t; NULL = the context's own stream */
} vkgv_frame_desc;

typedef struct vkgv_frame_stats {
    uint32_t draws;                  /* meshlet draws submitted */
    uint32_t visible_early;          /* draws rasterized in the first (or only) pass */
    uint32_t visible_late;           /* draws rasterized in the late pass (two-pass) */
    uint32_t big_triangles;          /* triangles sent to the tiled path */
    uint32_t big_tiles;              /* (32x32 screen tile, triangle) pairs they were binned to (frames
                                        with VKGV_STATS; 0 otherwise) */
    uint32_t clipped_triangles;      /* triangles that needed near/guard-band clipping */
    uint32_t overflow;               /* non-zero if a queue was full: bit 0 big-triangle queue, bit 1 clip
                                        queue, bit 2 big-triangle chunk list, bit 3 a screen tile's
                                        triangle list.  The frame is still complete: those triangles took
                                        the slower direct path (the wave sweeps the triangle, or its part
                                        in that tile) */
    uint32_t timed_frames;           /* frames averaged in the ms_* fields */
    /* per-pass GPU milliseconds (HIP events), averaged over the frames rendered
     * with VKGV_TIMING since the previous vkgv_view_read_stats (at most 64) */
    float    ms_clear, ms_cull, ms_raster, ms_hiz, ms_cull_late, ms_raster_late,
             ms_resolve, ms_pyramid, ms_total;
    uint32_t max_tile_tris;          /* most big triangles binned to one 32x32 tile (frames with
                                        VKGV_STATS; 0 otherwise) */
    /* raster counts of the last frame rendered with VKGV_STATS (0 otherwise) */
    uint64_t raster_triangles;       /* triangles past the back-face test set up for rasterization */
    uint64_t fragments;              /* covered pixel centres with 0 <= z <= 1: visibility-buffer updates */
    uint64_t fragment_atomics;       /* of those, written with a 64-bit atomicMin */
    uint64_t fragment_stores;        /* tile merges written with a plain read-min-store */
} vkgv_frame_stats;

/* Error-path testing: when the environment variable VKGV_FAULT_ALLOC_AFTER=n
 * is set at vkgv_create, that context's first n device allocations succeed
 * and every later one fails with VKGV_ERROR_OUT_OF_MEMORY. */
typedef struct vkgv_create_info {
    int32_t  device;                 /* HIP device ordinal */
    uint32_t big_triangle_capacity;  /* per view and raster pass; 0 = default (2^18), at most 2^23 */
    uint32_t clip_triangle_capacity; /* per view and raster pass; 0 = default (2^16), at most 2^22 */
    uint32_t tile_triangle_capacity; /* big triangles one 32x32 screen tile's list holds per raster pass;
                                        0 = default (min(1024, 2^24 / tiles), at least 64); tiles x capacity
                                        at most 2^28 */
} vkgv_create_info;

typedef struct vkgv_ctx  vkgv_ctx;
typedef struct vkgv_view vkgv_view;
typedef struct vkgv_batch vkgv_batch;

/* Raster tunables (the autotuner's knobs; per context). */
typedef struct vkgv_tuning {
    uint32_t small_max_extent;       /* pixels: triangles with a smaller bounding box are rasterized by
                                        the meshlet's own wave, larger ones are binned (1..64, default 16) */
    uint32_t raster_blocks_per_cu;   /* persistent meshlet-raster grid = CUs * this (1..16, default 3) */
    uint32_t tile_item_tris;         /* binned path: a tile whose list holds more triangles is split into
                                        up to 8 work items of about this many (1..256, default 8) */
    uint32_t resolve_cache_slots;    /* rasterized meshlets per view whose triangles are cached for the
                                        barycentric resolve (0..8192, default 8192); the resolve re-fetches
                                        and re-transforms the triangles of the others (same results) */
} vkgv_tuning;

/* Status codes */
#define VKGV_OK                       0
#define VKGV_ERROR_INVALID_ARGUMENT (-1)
#define VKGV_ERROR_OUT_OF_MEMORY    (-2)
#define VKGV_ERROR_HIP              (-3)
#define VKGV_ERROR_LIMIT            (-4)
#define VKGV_ERROR_STATE            (-5)

/* ------------------------------------------------------------------------ */
/* Context.  Replaces Device + World's GPU buffers (device.cpp:99-290,
 * world.cpp:10-185).                                                        */
int         vkgv_create(const vkgv_create_info* info, vkgv_ctx** out);
void        vkgv_destroy(vkgv_ctx* ctx);
const char* vkgv_last_error(const vkgv_ctx* ctx);
int         vkgv_sync(vkgv_ctx* ctx);      /* waits for the context's stream */
/* Raster tunables (no counterpart in the reference, whose raster is fixed
 * function): set by the autotuner, take effect at the next frame. */
int         vkgv_set_tuning(vkgv_ctx* ctx, const vkgv_tuning* tuning);
int         vkgv_get_tuning(const vkgv_ctx* ctx, vkgv_tuning* tuning);

/* The 255 ascending thresholds the resolve uses for fromLinear + unorm8
 * (srgb.h.glsl:26-32): the 8-bit code of a linear value c is the number of
 * thresholds <= c.  Host only; needs no device. */
void        vkgv_srgb_thresholds(float out[255]);

/* Calibration (no counterpart in the reference, whose raster is fixed
 * function): the device's rate of 64-bit atomicMin without a returned value,
 * the raster's fragment write, on a 3840 x 2160 key buffer the call allocates
 * and frees.  scattered = 0: the lanes of a wave update consecutive keys;
 * 1: pseudo-random keys.  Best of three timed launches; synchronous. */
int         vkgv_measure_atomic_rate(vkgv_ctx* ctx, uint32_t scattered, double* atomics_per_second);

/* Copies meshes, primitives and materials to device memory owned by the
 * context (assets.cpp:375-429 staging upload, world.cpp:10-185 primitive and
 * material buffers).  Replaces any previous scene. */
int vkgv_upload_scene(vkgv_ctx* ctx,
                      const vkgv_mesh_desc* meshes, uint32_t mesh_count,
                      const vkgv_primitive_desc* primitives, uint32_t primitive_count,
                      const vkgv_material* materials, uint32_t material_count);

/* vkgv_upload_scene with the mesh arrays in caller-owned device memory: the
 * pointers of each vkgv_mesh_desc (indices, vertices, meshlets, cones) are
 * device addresses, as the reference's Primitive holds buffer device
 * addresses (mesh_common.h.glsl:102-113, filled in world.cpp:10-185); the
 * descriptors, primitives and materials themselves are host arrays.  The
 * library copies device to device into its packed buffers and checks every
 * meshlet on the device (counts, ranges, vertex and triangle indices) before
 * accepting the scene; the caller's buffers may be freed on return. */
int vkgv_upload_scene_device(vkgv_ctx* ctx,
                             const vkgv_mesh_desc* meshes, uint32_t mesh_count,
                             const vkgv_primitive_desc* primitives, uint32_t primitive_count,
                             const vkgv_material* materials, uint32_t material_count);

/* Albedo textures of the textured resolve: the sampling the reference's
 * fragment shader leaves commented out (visbuffer.frag.glsl:31-32,
 * `color * albedoFactor * toLinear(texture(textures[albedoIndex],
 * transformUv(material, uv)))`), with transformUv of mesh_common.h.glsl:
 * 129-135.  Texels are sRGB-encoded RGBA8, decoded per texel (as an _SRGB
 * image format does) and filtered bilinearly at level 0 in linear space;
 * alpha is linear.  Wrap modes per axis as glTF samplers give them. */
#define VKGV_WRAP_REPEAT          0u   /* glTF 10497 */
#define VKGV_WRAP_CLAMP_TO_EDGE   1u   /* glTF 33071 */
#define VKGV_WRAP_MIRRORED_REPEAT 2u   /* glTF 33648 */
typedef struct vkgv_texture_desc {
    const uint8_t* rgba8;            /* width * height * 4 bytes, row-major, row 0 at v = 0 */
    uint32_t width, height;          /* 1 .. 16384 */
    uint32_t wrap_s, wrap_t;         /* VKGV_WRAP_* */
} vkgv_texture_desc;

/* Copies the textures to device memory owned by the context, replacing any
 * previous set (no counterpart in the reference, whose image load task is
 * empty: assets.cpp:467-471).  A frame with VKGV_RESOLVE_TEXTURED fails with
 * VKGV_ERROR_STATE while a material's albedo_index is neither ~0u nor below
 * `count`. */
int vkgv_upload_textures(vkgv_ctx* ctx, const vkgv_texture_desc* textures, uint32_t count);

/* The draw list (World::rebuildDrawBuffer, world.cpp:230-293): at most 2^25
 * draws, each index checked against the uploaded scene. */
int vkgv_set_draws(vkgv_ctx* ctx, const vkgv_meshlet_draw* draws, uint32_t count);

/* Per-node mat4 transforms (World::updateTransformBuffer, world.cpp:295-345).
 * The transforms they replace become the previous frame's (motion vectors). */
int vkgv_set_transforms(vkgv_ctx* ctx, const float* mat4s, uint32_t count);
/* The same from device memory (a transform buffer another kernel wrote, e.g.
 * a training step's parameters): copied device to device on `stream` (NULL:
 * the context's stream), with no host synchronization unless the context's
 * transform buffers must grow.  `mat4s` may be reused once the stream has
 * passed the call. */
int vkgv_set_transforms_device(vkgv_ctx* ctx, const float* mat4s, uint32_t count, void* stream);

/* ---- Node hierarchy and animations on the device ------------------------ */
/* glTF node (TRS form, as fastgltf's DecomposeNodeMatrices gives it). */
typedef struct vkgv_node {
    int32_t  mesh;                   /* glTF mesh index or -1 */
    uint32_t first_child, child_count;
    uint32_t pad;
    float    translation[3];
    float    rotation[4];            /* quaternion x, y, z, w */
    float    scale[3];
} vkgv_node;

/* AnimationSampler (assets.hpp:36-110): keyframe times times[input_offset ..
 * + input_count) (ascending), output values from values[output_offset]
 * (3 or 4 floats per output; cubic splines store in-tangent, value,
 * out-tangent per key). */
#define VKGV_INTERP_LINEAR       0u
#define VKGV_INTERP_STEP         1u
#define VKGV_INTERP_CUBICSPLINE  2u
typedef struct vkgv_anim_sampler {
    uint32_t input_offset, input_count;
    uint32_t output_offset;
    uint32_t interpolation;          /* VKGV_INTERP_* */
} vkgv_anim_sampler;

/* AnimationChannel: the node and TRS component a sampler drives.  Channels
 * are applied in array order (the reference's animation-then-channel order,
 * world.cpp:194-216): a later channel on the same node and path wins. */
#define VKGV_PATH_TRANSLATION    0u
#define VKGV_PATH_ROTATION       1u
#define VKGV_PATH_SCALE          2u
typedef struct vkgv_anim_channel {
    uint32_t node, sampler, path, pad;
} vkgv_anim_channel;

/* Uploads the node hierarchy of the scene's roots and its animations for
 * vkgv_update_transforms.  The mesh nodes, in the depth-first order of
 * World::iterateNode, own transforms 0 .. n-1 (the order vkgv_world_build
 * and the reference's draw list use), and the context's transform count
 * becomes n.  Node n's children are children[first_child .. + child_count),
 * which must lie within the `children_count` entries.  The nodes reached
 * from the roots must form a forest (no node reached twice), as glTF
 * requires; depth is limited to 64. */
int vkgv_set_hierarchy(vkgv_ctx* ctx, const vkgv_node* nodes, uint32_t node_count, const uint32_t* children,
                       uint32_t children_count, const uint32_t* roots, uint32_t root_count,
                       const vkgv_anim_sampler* samplers, uint32_t sampler_count,
                       const float* times, uint32_t time_count, const float* values, uint32_t value_count,
                       const vkgv_anim_channel* channels, uint32_t channel_count);

/* The per-frame transform update of World::updateDrawBuffers
 * (world.cpp:346-354) on the device: every animated node's TRS sampled at
 * `time` (AnimationSampler::sample, assets.hpp:52-110, looping with
 * fmod(time, last key)), then each mesh node's world matrix
 * parent * T * R * S down its ancestor chain (World::iterateNode,
 * world.cpp:187-228), written to the transform buffer the frames read.  The
 * transforms it replaces become the previous frame's.  Kernels and one
 * device copy only: no allocation, so it may be captured in a graph. */
int vkgv_update_transforms(vkgv_ctx* ctx, float time, void* stream);
/* Copies the current transforms to the host (float[count][16]). */
int vkgv_read_transforms(vkgv_ctx* ctx, float* mat4s, uint32_t count);

/* ------------------------------------------------------------------------ */
/* Views: the per-resolution images of the frame (visbuffer, depth, Hi-Z
 * pyramid, colour), application.cpp:181-576.                                */
int  vkgv_view_create(vkgv_ctx* ctx, uint32_t width, uint32_t height, vkgv_view** out);
void vkgv_view_destroy(vkgv_view* view);

/* Records one frame on the stream: cull + compact, rasterize to the 64-bit
 * visibility buffer, resolve, Hi-Z (application.cpp:843-1003). */
int vkgv_render(vkgv_ctx* ctx, vkgv_view* view, const vkgv_frame_desc* frame);

/* Builds the Hi-Z pyramid from the view's current depth
 * (application.cpp:951-1003 + hiz_reduce.comp.glsl). */
int vkgv_view_build_pyramid(vkgv_ctx* ctx, vkgv_view* view, void* stream);

/* Forgets the two-pass visibility history (every draw "not visible"). */
int vkgv_view_reset_history(vkgv_ctx* ctx, vkgv_view* view);

/* Pyramid geometry: levels = floor(log2(max(W,H))) (application.cpp:472-473);
 * level k is max(1, W >> (k+1)) x max(1, H >> (k+1)).  sizes gets 2*levels
 * values (w0,h0,w1,h1,...); texels is the total float count. */
int vkgv_view_pyramid_info(const vkgv_view* view, uint32_t* levels, uint32_t* sizes, uint64_t* texels);

/* Readbacks / uploads (synchronous on the context's stream). */
int vkgv_view_read_visbuffer(vkgv_ctx* ctx, vkgv_view* view, uint64_t* keys);      /* W*H */
int vkgv_view_read_ids_depth(vkgv_ctx* ctx, vkgv_view* view, uint32_t* ids, float* depth);
int vkgv_view_read_color(vkgv_ctx* ctx, vkgv_view* view, uint32_t* rgba8);         /* W*H */
/* Motion vectors of the last frame rendered with VKGV_MOTION_VECTORS,
 * float[H][W][2]: ((prev.xy / prev.w) * 0.5 - 0.5) - ((pos.xy / pos.w) * 0.5 - 0.5)
 * (visbuffer.frag.glsl:38), where prev = prevViewProjection * previous
 * transform * vertex, interpolated with the pixel's perspective-correct
 * barycentrics, and pos.xy / pos.w is the pixel centre's NDC; (0, 0) where
 * nothing was drawn. */
int vkgv_view_read_motion(vkgv_ctx* ctx, vkgv_view* view, float* motion);
int vkgv_view_read_pyramid(vkgv_ctx* ctx, vkgv_view* view, float* texels);
int vkgv_view_write_pyramid(vkgv_ctx* ctx, vkgv_view* view, const float* texels);
int vkgv_view_write_depth(vkgv_ctx* ctx, vkgv_view* view, const float* depth);     /* ids cleared */
/* pass 0: first/only pass list, pass 1: late pass list.  Order within a list
 * is unspecified (compare as sets); *count receives the full length. */
int vkgv_view_read_visible(vkgv_ctx* ctx, vkgv_view* view, uint32_t pass,
                           uint32_t* out, uint32_t capacity, uint32_t* count);
int vkgv_view_read_stats(vkgv_ctx* ctx, vkgv_view* view, vkgv_frame_stats* stats);

/* ------------------------------------------------------------------------ */
/* Batches: many views of uneven size rendered by one sequence of launches
 * (every kernel covers all views), optionally replayed as one captured
 * hipGraph (VKGV_GRAPH).  No counterpart in the reference, which records one
 * view per frame (application.cpp:843-1003); this is the batched form of the
 * same frame.  A batch refers to its views (at most 256, each once); destroy
 * the batch before its views.  After vkgv_render_batch the views' images and
 * readbacks are those of the batch frame. */
int  vkgv_batch_create(vkgv_ctx* ctx, vkgv_view* const* views, uint32_t count, vkgv_batch** out);
void vkgv_batch_destroy(vkgv_batch* batch);
/* frames[i] is view i's camera (and eye); flags and stream must be the same
 * for every view.  With VKGV_GRAPH the frame is captured on first use (and
 * again after the flags, stream, tuning or draw list change) and replayed. */
int  vkgv_render_batch(vkgv_ctx* ctx, vkgv_batch* batch, const vkgv_frame_desc* frames, uint32_t count);
/* Counters summed over the views; ms_total averaged over the VKGV_TIMING
 * frames since the previous call. */
int  vkgv_batch_read_stats(vkgv_ctx* ctx, vkgv_batch* batch, vkgv_frame_stats* stats);

/* Autotuner for C and C++ hosts (no counterpart in the reference, whose raster
 * is fixed function): renders the batch frame `frames` (as vkgv_render_batch
 * would, flags and stream from frames[0]) under candidate raster tunings —
 * small-triangle extent 4/8/16/32, raster blocks per CU 2/3, tile work item
 * 8/16/32 triangles, resolve cache 0/all slots — by a coordinate search (one
 * knob at a time, the others at the best so far, up to three sweeps), each
 * timed with HIP events over `frames_per_candidate` frames after one warm-up
 * frame; tunings whose frames fill a raster queue are rejected.  The fastest
 * is applied to the context and returned (with its ms per frame); the views
 * hold the last candidate's frame.  Synchronous. */
int  vkgv_autotune_batch(vkgv_ctx* ctx, vkgv_batch* batch, const vkgv_frame_desc* frames, uint32_t count,
                         uint32_t frames_per_candidate, vkgv_tuning* best, float* best_ms);

/* ------------------------------------------------------------------------ */
/* Differentiable frame: gradients of an image loss with respect to the
 * instance transforms, through analytic edge antialiasing of the view's last
 * frame.  No counterpart in the reference, which has no differentiable path;
 * the image is its flat resolve (visbuffer_resolve.comp.glsl:17-41) in linear
 * colour.  `camera` must be the one the view's last frame was rendered with,
 * and the scene and transforms those of that frame.  Device pointers; the
 * calls are stream-ordered (stream NULL: the context's stream).
 *
 * image: float [H][W][3], the albedo factor of each pixel's material
 * (background 0) with each pair of horizontally or vertically adjacent
 * pixels of different colour blended along the front triangle's crossed
 * edge (see csrc/grad.hip). */
int vkgv_view_aa_image(vkgv_ctx* ctx, vkgv_view* view, const vkgv_camera* camera, float* image, void* stream);
/* vkgv_view_aa_image without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under graph capture,
 * serving loops).  The same kernels and the same image, bit for bit, but the
 * per-pixel scratch is a fixed array in the library (one per GPU, held until
 * the process ends) instead of a buffer the view allocates on first use.
 * Views of more than 3840 x 2160 pixels fail with VKGV_ERROR_LIMIT.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (other
 * streams, other contexts of the device) race. */
int vkgv_view_aa_image_noalloc(vkgv_ctx* ctx, vkgv_view* view, const vkgv_camera* camera, float* image, void* stream);
/* g_transforms (double [transform_count][16], column-major like
 * vkgv_set_transforms) += d(sum g_image . image) / d transforms, for
 * g_image float [H][W][3].  transform_count must be at least the count last
 * given to vkgv_set_transforms.  The rasterization has no gradient. */
int vkgv_view_aa_backward(vkgv_ctx* ctx, vkgv_view* view, const vkgv_camera* camera, const float* g_image,
                          double* g_transforms, uint32_t transform_count, void* stream);

/* Visibility-buffer key layout of this implementation: one 64-bit word per
 * pixel, (~float_bits(depth) << 32) | (0xFFFFFFFE - id), min-reduced with
 * atomicMin.  Reversed-Z GREATER_OR_EQUAL (larger depth wins, ties go to
 * the later primitive) becomes a plain unsigned minimum; the clear value is
 * all ones (depth 0.0, id 0xFFFFFFFF). */
#define VKGV_KEY_CLEAR 0xFFFFFFFFFFFFFFFFull

#ifdef __cplusplus
}
#endif
#endif /* VKGV_H */
