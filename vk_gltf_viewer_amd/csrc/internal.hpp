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

// Device view of the uploaded scene.
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

// Per-view frame counters in device memory.
struct Counters {
    uint32_t vis_count[2];          // pass 0 / late-pass visible lists
    uint32_t raster_cursor[2];      // meshlet work cursors
    unsigned long long big;         // (triangles << 32) | tiles of the current raster pass
    uint32_t clipped;
    uint32_t overflow;
    uint32_t hiz_done;
    uint32_t big_total_tris;        // statistics over the frame
    uint32_t big_total_tiles;
    uint32_t pad[5];
};
static_assert(sizeof(Counters) == 64, "Counters layout");

struct ViewDev {
    uint32_t width, height;
    unsigned long long* keys;       // W*H visibility-buffer keys
    uint32_t* color;                // W*H RGBA8
    float* pyramid;
    PyramidDesc pyr;
    uint32_t* visible[2];           // compacted draw lists
    uint32_t* history;              // two-pass visibility bits, ceil(draws/64)*2 words
    Counters* counters;
    BigTri* big;
    uint32_t* big_tile_start;
    uint32_t big_capacity;
};

// Camera as kernels take it (a by-value copy of vkgv_camera).
struct CameraArg {
    vkgv_camera cam;
    float eye[3];
};

enum CullMode : uint32_t { CULL_SINGLE = 0, CULL_EARLY = 1, CULL_LATE = 2 };

// ---- launchers (stream-ordered, no host synchronization) -----------------
void launch_build_cull_records(const SceneDev& s, CullRec* out, hipStream_t st);
void launch_clear_keys(unsigned long long* keys, uint64_t n, hipStream_t st);
void launch_reset_counters(Counters* c, uint32_t what, hipStream_t st);
void launch_cull(const SceneDev& s, const ViewDev& v, const CameraArg& cam, uint32_t flags, CullMode mode,
                 uint32_t list, hipStream_t st);
void launch_raster(const SceneDev& s, const ViewDev& v, const Mat4& vp, uint32_t flags, uint32_t list,
                   uint32_t grid_blocks, uint32_t small_max_extent, hipStream_t st);
void launch_raster_big(const ViewDev& v, uint32_t grid_blocks, hipStream_t st);
void launch_hiz(const ViewDev& v, hipStream_t st);
void launch_resolve(const SceneDev& s, const ViewDev& v, const Mat4& vp, uint32_t flags, hipStream_t st);
void launch_write_depth_keys(unsigned long long* keys, const float* depth, uint64_t n, hipStream_t st);

// Tunables of the raster (set by the autotuner; defaults otherwise).
struct RasterTuning {
    uint32_t small_max_extent;      // triangles whose bbox extent is below this (pixels) stay in the meshlet wave
    uint32_t raster_blocks_per_cu;  // persistent raster grid = CUs * this
    uint32_t big_blocks;            // tiled-path grid size
};

} // namespace vkgv
