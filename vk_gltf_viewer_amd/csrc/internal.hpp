// Internal declarations shared by the HIP translation units of libvkgv_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"
#include "vkgv.h"

namespace vkgv {

// Per-draw cull record, built on the device from the contract arrays when the
// draw list changes (the reference rebuilds its draw buffer only then,
// world.cpp:230-293).  32 bytes, read with two 16-byte loads per lane so that
// a wave's 64 records are one contiguous 2 KiB span.
struct CullRec {
    float cx, cy, cz, ex;   // meshlet AABB centre, extent.x
    float ey, ez;           // extent.y, extent.z
    uint32_t transform;     // transform index
    uint32_t cone;          // index into the cone table, ~0u = none
};
static_assert(sizeof(CullRec) == 32, "CullRec layout");

// What the raster needs of one meshlet draw, built per draw when the draw
// list changes and copied into the visible lists by the cull, so that a wave
// reaches a meshlet's indices in one dependent load instead of four
// (draw -> primitive -> meshlet -> buffers).  32 bytes.
struct RasterRec {
    uint32_t draw;                  // draw index
    uint32_t transform;             // transform index
    uint32_t vidx;                  // first vertex index, elements into the packed vertex-index buffer
    uint32_t tri;                   // first triangle byte, into the packed triangle buffer
    uint32_t vtx;                   // first vertex of the mesh, byte offset into the packed vertex buffer
    uint32_t material;              // material index | double_sided << 31
    uint32_t counts;                // vertex_count | triangle_count << 8
    uint32_t pad;
};
static_assert(sizeof(RasterRec) == 32, "RasterRec layout");

// Device view of the uploaded scene.
// Animated TRS of one node for the device transform update.
struct NodeTrs { float t[3], q[4], s[3]; };

// Node hierarchy and animations on the device (vkgv_set_hierarchy).
struct AnimDev {
    const vkgv_node* nodes;
    uint32_t node_count;
    const uint32_t* chan_first;          // node_count + 1: the node's channels, in application order
    const vkgv_anim_channel* channels;
    const vkgv_anim_sampler* samplers;
    const float* times;
    const float* values;
    NodeTrs* trs;                        // per node, this update
    uint32_t mesh_count;                 // mesh nodes = transforms
    const uint32_t* chain_first;         // mesh_count + 1
    const uint32_t* chain_nodes;         // root-to-node chain of each mesh node
};

// One albedo texture: `width * height` decoded texels from texels + offset.
struct TexRec {
    uint64_t offset;
    uint32_t width, height;
    uint32_t wrap_s, wrap_t;
    uint32_t pad[2];
};
static_assert(sizeof(TexRec) == 32, "TexRec layout");

// What the textured resolve reads of a material, in one record: the uv
// transform (scale, offset, cos and sin of the rotation, evaluated once on
// the host) and its albedo texture (width 0: none).
struct MatTex {
    float uv_scale[2], uv_offset[2];
    float cos_r, sin_r;
    uint32_t width, height;
    uint64_t offset;                // first texel in the texel buffer
    uint32_t wrap_s, wrap_t;
};
static_assert(sizeof(MatTex) == 48, "MatTex layout");

struct SceneDev {
    const vkgv_meshlet_draw* draws;
    uint32_t draw_count;
    const float* transforms;
    uint32_t transform_count;
    const vkgv_primitive* prims;
    uint32_t prim_count;
    const vkgv_material* materials;
    uint32_t material_count;
    const uint32_t* prim_cone_base;      // per primitive: first index of its mesh's cones (~0u = none)
    const vkgv_meshlet_cone* cones;      // cones of all meshes, concatenated
    const CullRec* cull;                 // per draw
    const RasterRec* rrec;               // per draw
    // packed mesh buffers of the scene (the primitives' device addresses point into them)
    const uint32_t* vidx_base;
    const uint8_t* tri_base;
    const vkgv_vertex* vtx_base;
    // sRGB encode + unorm8 tables (SrgbTables)
    const void* srgb;
    // textured resolve: texture records, their texels (RGBA8 words), the
    // per-byte sRGB decode, and each material's uv transform and texture
    const TexRec* textures;
    uint32_t texture_count;
    const float4* texels;           // decoded: toLinear(r, g, b), a / 255
    const float* srgb_decode;
    const MatTex* mat_tex;
    // the transforms of the previous update (motion vectors)
    const float* prev_transforms;
};

// The 8-bit code of fromLinear(c) is the number of the 255 ascending
// thresholds <= c (the encode is monotonic).  Buckets of the float's top 16
// bits hold at most one threshold each: code = base[bucket] + (c >= thr[base]).
struct SrgbTables {
    float thr[256];             // 255 thresholds, thr[255] = +inf
    uint32_t key0;              // top 16 bits of thr[0]
    uint32_t nbuckets;          // buckets from key0 to the top 16 bits of 1.0f
    uint8_t base[2048];         // thresholds below each bucket
};

// Triangle handed to the tiled (large / clipped) path: snapped vertices,
// per-vertex depth, id and pixel bounding box.  48 bytes.
struct BigTri {
    int32_t x[3], y[3];
    float z[3];
    uint32_t id;
    uint32_t bb0;   // px0 | py0 << 16
    uint32_t bb1;   // px1 | py1 << 16
};
static_assert(sizeof(BigTri) == 48, "BigTri layout");

// Rasterization setup of a big triangle, computed once when it is binned.
// Edge k runs from vertex (k+1)%3 to (k+2)%3 of the oriented triangle;
// e0[k] is its edge function at the centre of the bounding box's first pixel
// minus the top-left bias, so a pixel is covered iff every value is >= 0.
struct BigSetup {
    long long e0[3];
    int ex[3], ey[3];
    float za, zb, z0;
    int x0, y0;             // vertex 0, 1/256 pixel
    uint32_t id;
    uint32_t bb0, bb1;      // px0 | py0 << 16, px1 | py1 << 16
};
static_assert(sizeof(BigSetup) == 80, "BigSetup layout");

// Data of one rasterized triangle kept for the resolve: with v_k the clip
// (x, y, w) of vertex k, r[k] = v_{k+1} x v_{k+2} (the barycentric of vertex
// k is proportional to r[k] . (xn, yn, 1)), and the packed RGBA8 colours.
struct TriCache {
    float r[3][3];
    uint32_t col[3];
    uint32_t uv[3];                 // float16 uv pairs of the vertices (textured frames only)
    uint32_t pad;
};
static_assert(sizeof(TriCache) == 64, "TriCache layout");

// Triangle waiting for near-plane / guard-band clipping (clip-space vertices).
struct ClipTri {
    float4 c[3];
    uint32_t id;
    uint32_t pad[3];
};
static_assert(sizeof(ClipTri) == 64, "ClipTri layout");

// Per-view frame counters in device memory.
struct Counters {
    unsigned long long big_chunks;  // 64-tile chunks of the pair list reserved in the current raster pass
    uint32_t vis_count[2];          // pass 0 / late-pass visible lists
    uint32_t big_count;             // big-triangle slots reserved in the current raster pass
    uint32_t redo_count;            // meshlets of the current raster pass in `redo`
    uint32_t clip_count;            // triangles queued for clipping in the current pass
    uint32_t reserved1;
    uint32_t clipped;               // statistics over the frame
    uint32_t overflow;              // slow path taken: bit 0 big queue, bit 1 clip queue, bit 2 chunk list, bit 3 a tile list full
    uint32_t reserved2;
    uint32_t big_total_tris;        // big triangles over the frame's passes
    uint32_t big_total_pairs;       // their (tile, triangle) pairs (frames with VKGV_STATS)
    uint32_t max_tile_tris;         // most big triangles binned to one tile (frames with VKGV_STATS)
    uint32_t reserved3;
    uint32_t pad;
    // raster statistics over the frame, counted only by frames with VKGV_STATS
    unsigned long long raster_tris;     // triangles past the back-face test that were set up for raster
    unsigned long long fragments;       // covered pixel centres with 0 <= z <= 1 (visibility-buffer updates)
    unsigned long long frag_atomics;    // global 64-bit atomicMin writes of those updates
    unsigned long long frag_stores;     // tile merges written with a plain read-min-store instead
    // deferred fragment runs of the current pass: staged meshlets << kJobCountBits | jobs reserved
    unsigned long long frag_jobs;
    unsigned long long pad2[3];
};
static_assert(sizeof(Counters) == 128, "Counters layout");

constexpr uint32_t kTileSplit = 8;      // most tile-pass parts sharing one screen tile's triangle list

// Deferred fragment runs.  A meshlet whose small triangles cover more than
// kJobKeep + kJobSlots / 2 pixel slots keeps the first kJobKeep for its own
// raster wave and hands the rest, kJobSlots at a time, to the bin pass: its
// triangle setups and slot prefix are staged in a FragStage record (an
// opaque copy of the raster's LDS arrays, 124 x 52 B + 128 x 4 B + total),
// and each job names the record and its first slot.
constexpr uint32_t kJobKeep = 512;
constexpr uint32_t kJobSlots = 256;
constexpr uint32_t kFragStageWords = 1744;   // 16-byte multiple
struct FragStage {
    uint32_t w[kFragStageWords];
};
static_assert(sizeof(FragStage) % 16 == 0, "FragStage records are copied 16 bytes at a time");
constexpr uint32_t kJobCountBits = 38;           // jobs field of Counters::frag_jobs (at most 2^25 meshlets x 140 jobs)
constexpr uint32_t kFragStageCapacity = 1024;    // staged meshlets per view and pass (7.1 MB)
constexpr uint32_t kFragJobCapacity = 1u << 15;  // their jobs per view and pass
struct FragJob {
    uint32_t stage;                 // FragStage record, ~0u: reserved but unused
    uint32_t first;                 // first pixel slot of the run
};
// Resolve block: 64 columns x (4 * kResolveRows) rows; each thread resolves
// kResolveRows pixels of its column, their loads issued together.
constexpr uint32_t kResolveRows = 4;

constexpr uint32_t kBinTile = 32;   // screen tile of the binned path, pixels

struct ViewDev {
    uint32_t width, height;
    unsigned long long* keys;       // W*H visibility-buffer keys of this frame
    unsigned long long* keys_next;  // the other key buffer: cleared by this frame's resolve for the next frame
    uint32_t* color;                // W*H RGBA8
    float* pyramid;
    PyramidDesc pyr;
    RasterRec* visible[2];          // compacted visible draws (first pass / late pass)
    uint32_t* redo;                 // list positions of the pass's meshlets that found a queue full
    uint32_t* history;              // two-pass visibility bits, ceil(draws/64)*2 words
    Counters* counters;             // this frame's
    Counters* counters_next;        // the next frame's, zeroed by this frame's resolve
    BigTri* big;                    // big triangles of the current raster pass
    uint32_t* big_chunk;            // first 64-tile chunk of each big triangle
    uint32_t* chunk_tri;            // big triangle of each 64-tile chunk
    uint32_t big_capacity;
    ClipTri* clipq;
    uint32_t clip_capacity;
    uint32_t tiles_x, tiles_y;      // kBinTile screen tiles
    uint32_t* bin_count;            // per pass and tile (pass p at p * tile_count_stride): triangles binned; zeroed at frame start
    uint32_t* bins;                 // per tile: a list of tile_capacity big-triangle indices
    uint32_t tile_capacity;         // triangles a tile's list holds (the rest take the direct path)
    uint32_t chunk_capacity;        // 64-tile chunks of big triangles per pass
    BigSetup* setups;               // per big triangle
    FragStage* frag_stage;          // deferred fragment runs: staged meshlets
    FragJob* frag_jobs;             // and their jobs
    uint32_t stage_capacity, job_capacity;
    // resolve cache: triangle t of the meshlet in visible-list slot s is
    // tri_cache[s * 128 + t]; draw_slot[d] is the slot of draw d this frame
    TriCache* tri_cache;
    float4* slot_albedo;            // albedo factor of each slot's material
    uint32_t* slot_mat;             // material index of each slot's meshlet
    uint32_t* draw_slot;
    uint32_t cache_slots;
    float* mvp;                     // per transform, this frame: VP M (16 floats, column-major), det(M) at [16]
    uint32_t mvp_capacity;          // transforms the table holds
    float* pmvp;                    // per transform: prevVP * previous M (motion vectors), 16 floats
    uint32_t pmvp_capacity;
    float2* motion;                 // W*H motion vectors (VKGV_MOTION_VECTORS)
    uint32_t hiz_tile_levels;       // pyramid levels produced by the tile pass of the Hi-Z build
    uint32_t hiz_tail_start;        // first level of the one-block tail; levels in between: hiz_level_kernel
    uint32_t hiz_tiles_x;           // 32x32 level-0 tiles per row
    uint32_t resolve_blocks_x;      // 64-pixel resolve blocks per row
};

// Stride of a view's per-pass tile counts: the tile count rounded up to 4, so
// that each pass's array starts 16-byte aligned.
__host__ __device__ __forceinline__ uint32_t tile_count_stride(const ViewDev& v)
{
    return (v.tiles_x * v.tiles_y + 3u) & ~3u;
}

// Camera as kernels take it (a by-value copy of vkgv_camera).
struct CameraArg {
    vkgv_camera cam;
    float eye[3];
};

// The views of one frame as kernels see them.  A single-view frame passes its
// view and camera by value (views == cams == nullptr); a batch keeps one
// descriptor and one camera per view in device memory.  Kernels over pixels find their view through a block map:
// entry = view | local block << 8 (at most 256 views per batch).
struct BatchDev {
    const ViewDev* views;           // [nviews], device memory; nullptr: the one view is view0
    const CameraArg* cams;          // [nviews] or nullptr
    CameraArg cam0;
    ViewDev view0;                  // a single-view frame's view, by value (no table load)
    uint32_t nviews;
    const uint32_t* map_clear;      // next-frame clear blocks (frames without a resolve)
    const uint32_t* map_hiz;        // Hi-Z tile blocks
    const uint32_t* map_resolve;    // resolve blocks
};

constexpr uint32_t kMaxBatchViews = 256;
constexpr uint32_t kClearKeysPerBlock = 256u * 32u;   // keys cleared by one frame-start block

__device__ __forceinline__ const CameraArg& batch_camera(const BatchDev& b, uint32_t v)
{
    return b.cams ? b.cams[v] : b.cam0;
}

__device__ __forceinline__ const ViewDev& batch_view(const BatchDev& b, uint32_t v)
{
    return b.views ? b.views[v] : b.view0;
}

// Host-side sizes of a batch's launches.
// Orders a wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kMvpStride = 20;          // floats per transform in a view's VP M table (80-byte rows)
constexpr uint32_t kHizMaxLevels = 16;
constexpr uint32_t kHizTailTexels = 8192;   // largest first tail level (the tail keeps it in LDS)

struct BatchGrid {
    uint32_t nviews;
    uint32_t clear_blocks, hiz_blocks, resolve_blocks;
    uint32_t max_tail_levels;       // any view needs the one-block Hi-Z tail
    uint32_t max_draws;
    uint32_t hiz_level_blocks[kHizMaxLevels];   // per level: 256-texel blocks of the largest view between tiles and tail
};

enum CullMode : uint32_t { CULL_SINGLE = 0, CULL_EARLY = 1, CULL_LATE = 2 };

// ---- launchers (stream-ordered, no host synchronization) -----------------
// Animation sampling and world matrices of the mesh nodes into `transforms`.
void launch_update_transforms(const AnimDev& d, float time, float* transforms, hipStream_t st);
void launch_build_cull_records(const SceneDev& s, CullRec* out, RasterRec* rrec, hipStream_t st);
void launch_clear_keys(unsigned long long* keys, uint64_t n, hipStream_t st);
// Frame start of every view of the batch: all counters zeroed, keys cleared,
// the view's table of VP M and det(M) per transform computed.
void launch_next_frame_clear(const BatchDev& b, const BatchGrid& g, hipStream_t st);
// Every view's prevViewProjection * previous transform table (motion vectors).
void launch_prev_mvp(const SceneDev& s, const BatchDev& b, const BatchGrid& g, hipStream_t st);
// Persistent cull + compaction of every view's draws; `cus` sizes the grid.
void launch_cull(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, CullMode mode,
                 uint32_t list, uint32_t cus, hipStream_t st);
void launch_raster(const SceneDev& s, const BatchDev& b, uint32_t flags, uint32_t list, uint32_t grid_blocks,
                   uint32_t small_max_extent, uint32_t cache_slots, hipStream_t st);
// Clipping, binning and tiled rasterization of the views' big-triangle queues.
void launch_raster_big(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, uint32_t list,
                       uint32_t small_max_extent, uint32_t tile_grid, uint32_t tile_item_tris,
                       hipStream_t st);
void launch_hiz(const BatchDev& b, const BatchGrid& g, hipStream_t st);
void launch_resolve(const SceneDev& s, const BatchDev& b, const BatchGrid& g, uint32_t flags, hipStream_t st);
void launch_write_depth_keys(unsigned long long* keys, const float* depth, uint64_t n, hipStream_t st);
// Differentiable frame (grad.hip): antialiased linear image of a view's keys,
// and its backward to the instance transforms (accumulated into g_transforms).
// mat: W*H scratch of per-pixel material indices (written, then read).
void launch_aa_image(const SceneDev& s, const unsigned long long* keys, uint32_t* mat, uint32_t W, uint32_t H,
                     const float* vp, float* image, hipStream_t st);
void launch_aa_backward(const SceneDev& s, const unsigned long long* keys, uint32_t* mat, uint32_t W, uint32_t H,
                        const float* vp, const float* g_image, double* g_transforms, hipStream_t st);
// One mesh of a scene uploaded from device memory, as its packed copies hold
// it: what the device-side check of its meshlets needs.
struct MeshCheck {
    uint64_t vidx_off;              // first vertex index, elements into the packed vertex-index buffer
    uint64_t vidx_count;
    uint64_t tri_off;               // first triangle byte, into the packed triangle buffer
    uint64_t tri_bytes;
    uint64_t vtx_count;
    uint64_t ml_off;                // first meshlet, bytes into the packed meshlet buffer
    uint64_t ml_first;              // meshlets of the meshes before this one
    uint32_t ml_count;
    uint32_t pad;
};
// Checks every meshlet of the packed scene (counts, ranges, vertex indices,
// triangle indices): *bad receives the smallest (mesh << 32 | meshlet) of a
// malformed one, and keeps its value (~0) when all are well formed.
hipError_t launch_check_meshes(const MeshCheck* meshes, uint32_t mesh_count, uint64_t meshlets, const uint32_t* vidx,
                               const uint8_t* tri, const uint8_t* ml, unsigned long long* bad, hipStream_t st);

// Calibration (calib.hip): 64-bit atomicMin throughput on a 3840 x 2160 key
// buffer of calib_key_count() keys, coalesced or scattered; 0 on success.
int measure_atomic_rate(unsigned long long* keys, uint32_t cus, uint32_t scattered, hipStream_t st, double* rate);
uint64_t calib_key_count();
uint32_t hiz_tile_levels(const PyramidDesc& pd);
uint32_t hiz_tail_start(const PyramidDesc& pd, uint32_t tile_levels);

// Tunables of the raster (set by the autotuner; defaults otherwise).
struct RasterTuning {
    uint32_t small_max_extent;      // triangles whose bbox extent is below this (pixels) stay in the meshlet wave
    uint32_t raster_blocks_per_cu;  // persistent raster grid = CUs * this
    uint32_t tile_item_tris;        // triangles per tile-pass work item before a tile is split
    uint32_t resolve_cache_slots;   // meshlets per view cached for the barycentric resolve
};

} // namespace vkgv
