// C ABI of libvkgv_hip.so (include/vkgv.h): context, scene upload, views,
// frame orchestration and readbacks.
//
// The frame (application.cpp:843-1003) becomes a sequence of stream-ordered
// launches with no host synchronization:
//   single pass:  reset, clear, cull (vs. previous pyramid), raster, tiled
//                 raster, resolve, [pyramid for the next frame]
//   two pass:     reset, clear, cull early (visible last frame), raster,
//                 pyramid, cull late (vs. that pyramid), raster, resolve
// Every index that a kernel dereferences is validated here on upload, so the
// kernels need no bounds checks on scene data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace vkgv;

namespace vkgv {
uint32_t hiz_tile_levels(const PyramidDesc& pd);
}

// Scratch of vkgv_view_aa_image_noalloc: the per-pixel material indices of
// the image pass, for views of at most kNoAllocPixels pixels (4K UHD, 33 MB).
// Each GPU the library runs on holds its own copy until the process ends.
// The pass's first kernel writes all W*H entries it later reads, so nothing
// needs clearing between calls.  Host code never uses the array's name as an
// address: vkgv_create resolves the device address with hipGetSymbolAddress
// (per context, with the context's device current).
constexpr size_t kNoAllocPixels = 3840u * 2160u;
static __device__ __align__(256) uint32_t scratch[kNoAllocPixels];

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

enum Ev { EV_START, EV_CLEAR, EV_CULL, EV_RASTER, EV_HIZ, EV_CULL_LATE, EV_RASTER_LATE, EV_RESOLVE, EV_PYRAMID, EV_COUNT };

} // namespace

struct vkgv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cus = 256;
    // scene
    std::vector<DevBuf> mesh_bufs;
    std::vector<uint32_t> mesh_meshlets;
    DevBuf prims, materials, cone_base, cones;
    uint32_t prim_count = 0, material_count = 0;
    std::vector<uint32_t> prim_meshlets;     // meshlet count per primitive (host copy for validation)
    bool scene = false;
    // frame inputs
    DevBuf draws, cull, rrec, transforms;
    uint32_t draw_count = 0, transform_count = 0;
    uint32_t max_transform = 0;
    bool draws_set = false;
    RasterTuning tuning{ 16u, 3u, 8u, 8192u };
    DevBuf srgb;                     // 255 sRGB-encode thresholds
    DevBuf srgb_decode;              // toLinear of each byte / 255 (256 floats)
    DevBuf tex_recs, texels, mat_tex;
    std::vector<vkgv_material> host_materials;   // what the per-material texture records are built from
    std::vector<TexRec> host_textures;
    uint32_t texture_count = 0;
    DevBuf prev_transforms;          // the transforms the last update replaced (motion vectors)
    uint32_t prev_count = 0;
    // node hierarchy and animations (vkgv_set_hierarchy)
    DevBuf an_nodes, an_chan_first, an_channels, an_samplers, an_times, an_values, an_trs, an_chain_first, an_chain;
    AnimDev anim{};
    bool anim_set = false;
    int64_t max_albedo_index = -1;   // largest texture handle a material names (-1: none)
    uint32_t big_capacity = 1u << 18;
    uint32_t clip_capacity = 1u << 16;
    uint32_t tile_capacity = 0;      // per-tile big-triangle list; 0: by view size
    uint32_t cache_slots = 8192;     // resolve-cache meshlet slots per view
    uint64_t generation = 0;         // bumped when view tables must be rebuilt (draw lists reallocated)
    // Allocation fault injection for error-path tests: when VKGV_FAULT_ALLOC_AFTER=n
    // is set at vkgv_create, the context's first n device allocations succeed and
    // every later one fails with VKGV_ERROR_OUT_OF_MEMORY (-1: off).
    int64_t fault_alloc_after = -1;
    // device address of `scratch` on this context's device
    uint32_t* noalloc_scratch = nullptr;
};

struct vkgv_view {
    vkgv_ctx* ctx = nullptr;
    ViewDev dev{};
    DevBuf keys, color, pyramid, visible0, visible1, history, counters, big, clipq, bin_count, bins;
    DevBuf keys_b, counters_b;       // the second key buffer and counters (frames alternate between them)
    DevBuf setups, tri_cache, slot_albedo, slot_mat, draw_slot, big_chunk, chunk_tri, redo;
    DevBuf frag_stage, frag_jobs;    // deferred fragment runs of heavy meshlets
    DevBuf aa_mat;                   // per-pixel material indices of the differentiable frame (allocated on first use)
    DevBuf mvp_tab;                  // per-frame VP M table, kMvpStride floats per transform
    DevBuf pmvp_tab, motion;         // prevVP * previous M table and W*H motion vectors (allocated on first use)
    uint32_t list_capacity = 0;
    // Ring of per-frame event sets: timed frames are averaged at the next
    // read_stats without synchronizing between frames.
    static constexpr uint32_t kRing = 64;
    hipEvent_t ev[kRing][EV_COUNT] = {};
    uint32_t ring_next = 0, ring_frames = 0;
    uint32_t ring_flags[kRing] = {};
    bool timed = false;
    uint32_t last_flags = 0;
    hipStream_t last_stream = nullptr;
    // the view as a one-view batch: block maps and launch sizes
    DevBuf maps;
    BatchDev batch{};
    BatchGrid grid{};
};

// A set of views rendered by one sequence of launches (optionally one
// captured hipGraph).
struct vkgv_batch {
    vkgv_ctx* ctx = nullptr;
    std::vector<vkgv_view*> views;
    DevBuf maps, cams;
    // Two device tables of the views' descriptors (and a captured frame for
    // each): frames alternate between the views' two key buffers, so in a run
    // of batch frames the tables alternate too.  A table is identified by the
    // key buffer each view renders into and the context generation.
    static constexpr uint32_t kTables = 2;
    DevBuf table[kTables];
    std::vector<const void*> table_keys[kTables];
    uint64_t table_generation[kTables] = { ~0ull, ~0ull };
    uint32_t table_next = 0;
    BatchDev batch{};
    BatchGrid grid{};
    // pinned staging ring for the per-frame cameras
    static constexpr uint32_t kStage = 4;
    CameraArg* staging[kStage] = {};
    hipEvent_t staged[kStage] = {};
    bool stage_used[kStage] = {};
    uint32_t stage_next = 0;
    // captured frame, per table
    hipGraphExec_t exec[kTables] = {};
    uint32_t exec_flags[kTables] = {};
    hipStream_t exec_stream[kTables] = {};
    RasterTuning exec_tuning[kTables] = {};
    // whole-frame timing ring
    static constexpr uint32_t kRing = 64;
    hipEvent_t ev0[kRing] = {}, ev1[kRing] = {};
    uint32_t ring_next = 0, ring_frames = 0;
    hipStream_t last_stream = nullptr;
};

namespace {

int fail(vkgv_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail((ctx), VKGV_ERROR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));   \
    } while (0)

void dfree(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int dalloc(vkgv_ctx* c, DevBuf& b, size_t bytes)
{
    dfree(b);
    if (bytes == 0) bytes = 16;
    if (c && c->fault_alloc_after >= 0) {
        if (c->fault_alloc_after == 0)
            return fail(c, VKGV_ERROR_OUT_OF_MEMORY, "hipMalloc: out of memory (injected by VKGV_FAULT_ALLOC_AFTER)");
        c->fault_alloc_after--;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, e == hipErrorOutOfMemory ? VKGV_ERROR_OUT_OF_MEMORY : VKGV_ERROR_HIP,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return VKGV_OK;
}

int upload(vkgv_ctx* c, DevBuf& b, const void* src, size_t bytes)
{
    int rc = dalloc(c, b, bytes);
    if (rc) return rc;
    if (bytes) HIP_TRY(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    return VKGV_OK;
}

// The per-material records of the textured resolve, from the materials and
// textures uploaded so far (a material whose texture is not uploaded yet gets
// none; frames that need it are refused until it is).  cos and sin of each uv
// rotation are evaluated once here, so that every pixel uses the same matrix.
static int rebuild_mat_tex(vkgv_ctx* c)
{
    std::vector<MatTex> mt(std::max<size_t>(c->host_materials.size(), 1));
    for (size_t m = 0; m < c->host_materials.size(); ++m) {
        const vkgv_material& mat = c->host_materials[m];
        MatTex& r = mt[m];
        r.uv_scale[0] = mat.uv_scale[0];
        r.uv_scale[1] = mat.uv_scale[1];
        r.uv_offset[0] = mat.uv_offset[0];
        r.uv_offset[1] = mat.uv_offset[1];
        r.cos_r = std::cos(mat.uv_rotation);
        r.sin_r = std::sin(mat.uv_rotation);
        const uint32_t ti = mat.albedo_index;
        if (ti != 0xFFFFFFFFu && ti < c->host_textures.size()) {
            const TexRec& t = c->host_textures[ti];
            r.width = t.width;
            r.height = t.height;
            r.offset = t.offset;
            r.wrap_s = t.wrap_s;
            r.wrap_t = t.wrap_t;
        }
    }
    int rc = upload(c, c->mat_tex, mt.data(), sizeof(MatTex) * mt.size());
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));   // mt is released on return
    return VKGV_OK;
}

SceneDev scene_dev(const vkgv_ctx* c)
{
    SceneDev s{};
    s.draws = static_cast<const vkgv_meshlet_draw*>(c->draws.p);
    s.draw_count = c->draw_count;
    s.transforms = static_cast<const float*>(c->transforms.p);
    s.transform_count = c->transform_count;
    s.prims = static_cast<const vkgv_primitive*>(c->prims.p);
    s.prim_count = c->prim_count;
    s.materials = static_cast<const vkgv_material*>(c->materials.p);
    s.material_count = c->material_count;
    s.prim_cone_base = static_cast<const uint32_t*>(c->cone_base.p);
    s.cones = static_cast<const vkgv_meshlet_cone*>(c->cones.p);
    s.cull = static_cast<const CullRec*>(c->cull.p);
    s.rrec = static_cast<const RasterRec*>(c->rrec.p);
    if (c->mesh_bufs.size() == 4) {
        s.vidx_base = static_cast<const uint32_t*>(c->mesh_bufs[0].p);
        s.tri_base = static_cast<const uint8_t*>(c->mesh_bufs[1].p);
        s.vtx_base = static_cast<const vkgv_vertex*>(c->mesh_bufs[2].p);
    }
    s.srgb = c->srgb.p;
    s.textures = static_cast<const TexRec*>(c->tex_recs.p);
    s.texture_count = c->texture_count;
    s.texels = static_cast<const float4*>(c->texels.p);
    s.srgb_decode = static_cast<const float*>(c->srgb_decode.p);
    s.mat_tex = static_cast<const MatTex*>(c->mat_tex.p);
    s.prev_transforms = static_cast<const float*>(c->prev_transforms.p ? c->prev_transforms.p : c->transforms.p);
    return s;
}

// toLinear (srgb.h.glsl:17-23) of every byte value b / 255: the per-texel
// decode of the textured resolve, evaluated once on the host.
void srgb_decode_table(float* out)
{
    for (int b = 0; b < 256; ++b) {
        const float v = (float)b / 255.0f;
        const bool cutoff = v < 0.04045f;
        const float higher = std::pow((v + 0.055f) / 1.055f, 2.4f);
        const float lower = v / 12.92f;
        out[b] = cutoff ? lower : higher;
    }
}

// fromLinear (srgb.h.glsl:9-15) followed by the unorm8 store.
uint32_t srgb_code_ref(float c)
{
    const bool cutoff = c < 0.0031308f;
    const float higher = 1.055f * std::pow(c, 1.0f / 2.4f) - 0.055f;
    const float lower = c * 12.92f;
    float v = cutoff ? lower : higher;
    v = 1.0f < v ? 1.0f : v;
    v = v < 0.0f ? 0.0f : v;
    if (!(v >= 0.0f)) v = 0.0f;
    return (uint32_t)(v * 255.0f + 0.5f);
}

// Thresholds t[k-1] = the smallest float c in [0, 1] whose code is >= k,
// found by bisection over the bit patterns (the code is non-decreasing in c).
void srgb_thresholds(float* t)
{
    for (uint32_t k = 1; k <= 255; ++k) {
        uint32_t lo = 0, hi = 0x3F800000u;   // code(lo) < k <= code(hi)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float f;
            std::memcpy(&f, &mid, 4);
            if (srgb_code_ref(f) >= k) hi = mid; else lo = mid;
        }
        std::memcpy(&t[k - 1], &hi, 4);
    }
}

Mat4 to_mat4(const float* m)
{
    Mat4 r;
    std::memcpy(r.m, m, sizeof r.m);
    return r;
}

int ensure_mvp(vkgv_ctx* c, vkgv_view* v)
{
    if (v->dev.mvp_capacity >= c->transform_count && v->mvp_tab.p) return VKGV_OK;
    const uint32_t cap = std::max<uint32_t>(c->transform_count, 64u);
    HIP_TRY(c, hipDeviceSynchronize());   // frames of other streams may still read the old table
    int rc;
    if ((rc = dalloc(c, v->mvp_tab, sizeof(float) * kMvpStride * (size_t)cap))) return rc;
    v->dev.mvp = static_cast<float*>(v->mvp_tab.p);
    v->dev.mvp_capacity = cap;
    c->generation++;
    return VKGV_OK;
}

// Motion-vector buffers of a view: the prevVP * previous M table and the image.
int ensure_motion(vkgv_ctx* c, vkgv_view* v)
{
    if (v->dev.pmvp_capacity >= c->transform_count && v->pmvp_tab.p && v->motion.p) return VKGV_OK;
    HIP_TRY(c, hipDeviceSynchronize());   // frames of other streams may still read the old buffers
    const uint32_t cap = std::max<uint32_t>(c->transform_count, 64u);
    int rc;
    if ((rc = dalloc(c, v->pmvp_tab, sizeof(float) * 16 * (size_t)cap))) return rc;
    v->dev.pmvp = static_cast<float*>(v->pmvp_tab.p);
    v->dev.pmvp_capacity = cap;
    if (!v->motion.p) {
        if ((rc = dalloc(c, v->motion, sizeof(float2) * (size_t)v->dev.width * v->dev.height))) return rc;
        HIP_TRY(c, hipMemsetAsync(v->motion.p, 0, v->motion.bytes, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        v->dev.motion = static_cast<float2*>(v->motion.p);
    }
    c->generation++;
    return VKGV_OK;
}

int ensure_lists(vkgv_ctx* c, vkgv_view* v)
{
    int rc0 = ensure_mvp(c, v);
    if (rc0) return rc0;
    if (v->list_capacity >= c->draw_count && v->visible0.p) return VKGV_OK;
    const uint32_t cap = std::max<uint32_t>(c->draw_count, 64u);
    int rc;
    if ((rc = dalloc(c, v->visible0, sizeof(RasterRec) * cap))) return rc;
    if ((rc = dalloc(c, v->visible1, sizeof(RasterRec) * cap))) return rc;
    const size_t hwords = (cap + 63u) / 64u;
    if ((rc = dalloc(c, v->history, sizeof(uint64_t) * hwords))) return rc;
    if ((rc = dalloc(c, v->draw_slot, sizeof(uint32_t) * cap))) return rc;
    if ((rc = dalloc(c, v->redo, sizeof(uint32_t) * cap))) return rc;
    HIP_TRY(c, hipMemsetAsync(v->history.p, 0, v->history.bytes, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));   // frames may run on another stream
    v->list_capacity = cap;
    c->generation++;
    v->dev.visible[0] = static_cast<RasterRec*>(v->visible0.p);
    v->dev.visible[1] = static_cast<RasterRec*>(v->visible1.p);
    v->dev.history = static_cast<uint32_t*>(v->history.p);
    v->dev.draw_slot = static_cast<uint32_t*>(v->draw_slot.p);
    v->dev.redo = static_cast<uint32_t*>(v->redo.p);
    return VKGV_OK;
}

int check_view(vkgv_ctx* c, vkgv_view* v)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!v || v->ctx != c) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "view does not belong to this context");
    return VKGV_OK;
}

int sync_view(vkgv_ctx* c, vkgv_view* v)
{
    if (v->last_stream && v->last_stream != c->stream) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

// Block maps of the pixel-parallel launches of a set of views (entry = view |
// local block << 8) and the launch sizes.
int build_maps(vkgv_ctx* c, const std::vector<const ViewDev*>& views, DevBuf& maps, BatchDev& bd, BatchGrid& g)
{
    std::vector<uint32_t> clear, hiz, resolve;
    g = BatchGrid{};
    g.nviews = (uint32_t)views.size();
    for (uint32_t vi = 0; vi < views.size(); ++vi) {
        const ViewDev& v = *views[vi];
        const uint64_t npx = (uint64_t)v.width * v.height;
        const uint64_t nclear = (npx + kClearKeysPerBlock - 1) / kClearKeysPerBlock;
        const uint64_t nhiz = v.pyr.levels ? (uint64_t)((v.pyr.w[0] + 31) / 32) * ((v.pyr.h[0] + 31) / 32) : 0;
        const uint64_t nres = (uint64_t)v.resolve_blocks_x * ((v.height + 4 * kResolveRows - 1) / (4 * kResolveRows));
        if (std::max({ nclear, nhiz, nres }) >= (1ull << 24))
            return fail(c, VKGV_ERROR_LIMIT, "view too large for a batch block map");
        for (uint64_t k = 0; k < nclear; ++k) clear.push_back(vi | (uint32_t)(k << 8));
        for (uint64_t k = 0; k < nhiz; ++k) hiz.push_back(vi | (uint32_t)(k << 8));
        for (uint64_t k = 0; k < nres; ++k) resolve.push_back(vi | (uint32_t)(k << 8));
        if (v.hiz_tail_start < v.pyr.levels) g.max_tail_levels = std::max(g.max_tail_levels, v.pyr.levels);
        for (uint32_t k = v.hiz_tile_levels; k < v.hiz_tail_start && k < kHizMaxLevels; ++k)
            g.hiz_level_blocks[k] = std::max(g.hiz_level_blocks[k], (uint32_t)(((uint64_t)v.pyr.w[k] * v.pyr.h[k] + 255) / 256));
    }
    g.clear_blocks = (uint32_t)clear.size();
    g.hiz_blocks = (uint32_t)hiz.size();
    g.resolve_blocks = (uint32_t)resolve.size();
    std::vector<uint32_t> all;
    all.insert(all.end(), clear.begin(), clear.end());
    all.insert(all.end(), hiz.begin(), hiz.end());
    all.insert(all.end(), resolve.begin(), resolve.end());
    int rc = dalloc(c, maps, sizeof(uint32_t) * std::max<size_t>(all.size(), 1));
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(maps.p, all.data(), sizeof(uint32_t) * all.size(), hipMemcpyHostToDevice));
    const uint32_t* base = static_cast<const uint32_t*>(maps.p);
    bd.map_clear = base;
    bd.map_hiz = base + clear.size();
    bd.map_resolve = base + clear.size() + hiz.size();
    bd.nviews = g.nviews;
    return VKGV_OK;
}

// The view as the next frame uses it: frames alternate between two key
// buffers and two counter sets; the frame renders into the one the previous
// frame's resolve cleared and clears the other for the frame after.
ViewDev next_frame_view(const ViewDev& d)
{
    ViewDev n = d;
    std::swap(n.keys, n.keys_next);
    std::swap(n.counters, n.counters_next);
    return n;
}

// Device copy of the views' descriptors (synchronous; only when they change).
int upload_table(vkgv_ctx* c, const std::vector<const ViewDev*>& views, DevBuf& table, BatchDev& bd)
{
    std::vector<ViewDev> host;
    for (const ViewDev* v : views) host.push_back(*v);
    if (!table.p || table.bytes < sizeof(ViewDev) * host.size()) {
        int rc = dalloc(c, table, sizeof(ViewDev) * host.size());
        if (rc) return rc;
    }
    HIP_TRY(c, hipMemcpyAsync(table.p, host.data(), sizeof(ViewDev) * host.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    bd.views = static_cast<const ViewDev*>(table.p);
    return VKGV_OK;
}

// The frame's launches for a set of views (application.cpp:843-1003); `mark`
// records the pass boundaries when timing.
template <typename Mark>
void frame_launches(vkgv_ctx* c, const BatchDev& b, const BatchGrid& g, uint32_t flags, hipStream_t st, Mark mark)
{
    const SceneDev s = scene_dev(c);
    const bool two_pass = flags & VKGV_CULL_TWO_PASS;
    const uint32_t raster_grid = (uint32_t)c->cus * c->tuning.raster_blocks_per_cu;
    const uint32_t tile_grid = (uint32_t)c->cus * 8u;
    mark(EV_START);
    if (flags & VKGV_MOTION_VECTORS) launch_prev_mvp(s, b, g, st);
    mark(EV_CLEAR);
    launch_cull(s, b, g, flags, two_pass ? CULL_EARLY : CULL_SINGLE, 0, (uint32_t)c->cus, st);
    mark(EV_CULL);
    launch_raster(s, b, flags, 0, raster_grid, c->tuning.small_max_extent, c->tuning.resolve_cache_slots, st);
    launch_raster_big(s, b, g, flags, 0, c->tuning.small_max_extent, tile_grid, c->tuning.tile_item_tris, st);
    mark(EV_RASTER);
    if (two_pass) {
        launch_hiz(b, g, st);
        mark(EV_HIZ);
        launch_cull(s, b, g, flags, CULL_LATE, 1, (uint32_t)c->cus, st);
        mark(EV_CULL_LATE);
        launch_raster(s, b, flags, 1, raster_grid, c->tuning.small_max_extent, c->tuning.resolve_cache_slots, st);
        launch_raster_big(s, b, g, flags, 1, c->tuning.small_max_extent, tile_grid, c->tuning.tile_item_tris, st);
        mark(EV_RASTER_LATE);
    }
    // the resolve also clears the next frame's key buffer and zeroes its
    // counters; a frame without one does that here
    if (!(flags & VKGV_SKIP_RESOLVE)) launch_resolve(s, b, g, flags, st);
    else launch_next_frame_clear(b, g, st);
    mark(EV_RESOLVE);
    if (!two_pass && (flags & VKGV_BUILD_PYRAMID)) launch_hiz(b, g, st);
    mark(EV_PYRAMID);
}

int check_frame_inputs(vkgv_ctx* c, uint32_t flags)
{
    if (!c->draws_set) return fail(c, VKGV_ERROR_STATE, "render: no draw list");
    if (c->draw_count && c->max_transform >= c->transform_count)
        return fail(c, VKGV_ERROR_STATE, "render: draw references transform " + std::to_string(c->max_transform)
                                             + " but only " + std::to_string(c->transform_count) + " are set");
    if ((flags & VKGV_CULL_TWO_PASS) && (flags & VKGV_CULL_OCCLUSION_PREV))
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "render: CULL_TWO_PASS and CULL_OCCLUSION_PREV are exclusive");
    if ((flags & VKGV_RESOLVE_TEXTURED) && c->max_albedo_index >= (int64_t)c->texture_count)
        return fail(c, VKGV_ERROR_STATE, "render: a material names texture " + std::to_string(c->max_albedo_index)
                                             + " but " + std::to_string(c->texture_count) + " are uploaded");
    return VKGV_OK;
}

} // namespace

// ---------------------------------------------------------------------------
extern "C" {

int vkgv_create(const vkgv_create_info* info, vkgv_ctx** out)
{
    if (!out) return VKGV_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* c = new vkgv_ctx();
    c->device = info ? info->device : 0;
    if (const char* fa = std::getenv("VKGV_FAULT_ALLOC_AFTER")) {
        char* end = nullptr;
        const long long n = std::strtoll(fa, &end, 10);
        if (end != fa && n >= 0) c->fault_alloc_after = n;
    }
    if (info && info->big_triangle_capacity) c->big_capacity = info->big_triangle_capacity;
    if (info && info->clip_triangle_capacity) c->clip_capacity = info->clip_triangle_capacity;
    if (info && info->tile_triangle_capacity) c->tile_capacity = info->tile_triangle_capacity;
    if (c->big_capacity > (1u << 23) || c->clip_capacity > (1u << 22) || c->tile_capacity > (1u << 28)) {
        delete c;
        return VKGV_ERROR_LIMIT;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= c->device || c->device < 0) {
        delete c;
        return VKGV_ERROR_HIP;
    }
    if (hipSetDevice(c->device) != hipSuccess) { delete c; return VKGV_ERROR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return VKGV_ERROR_HIP; }
    {
        // this device's copy of the no-alloc scratch (the device is current)
        void* sp = nullptr;
        if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
            vkgv_destroy(c);
            return VKGV_ERROR_HIP;
        }
        c->noalloc_scratch = static_cast<uint32_t*>(sp);
    }
    SrgbTables tab;
    std::memset(&tab, 0, sizeof tab);
    srgb_thresholds(tab.thr);
    const uint32_t inf = 0x7F800000u;
    std::memcpy(&tab.thr[255], &inf, 4);
    uint32_t b0;
    std::memcpy(&b0, &tab.thr[0], 4);
    tab.key0 = b0 >> 16;
    tab.nbuckets = (0x3F800000u >> 16) - tab.key0 + 1u;
    bool ok = tab.nbuckets <= sizeof tab.base;
    for (uint32_t b = 0, k = 0; ok && b < tab.nbuckets; ++b) {
        // thresholds below bucket b, and at most one inside it
        uint32_t kb;
        while (k < 255 && (std::memcpy(&kb, &tab.thr[k], 4), (kb >> 16) < tab.key0 + b)) ++k;
        tab.base[b] = (uint8_t)k;
        uint32_t k2 = k;
        while (k2 < 255 && (std::memcpy(&kb, &tab.thr[k2], 4), (kb >> 16) == tab.key0 + b)) ++k2;
        ok = k2 - k <= 1;
    }
    float dec[256];
    srgb_decode_table(dec);
    if (!ok || upload(c, c->srgb, &tab, sizeof tab) != VKGV_OK || upload(c, c->srgb_decode, dec, sizeof dec) != VKGV_OK
        || hipStreamSynchronize(c->stream) != hipSuccess) {
        vkgv_destroy(c);
        return VKGV_ERROR_HIP;
    }
    *out = c;
    return VKGV_OK;
}

void vkgv_srgb_thresholds(float out[255])
{
    if (out) srgb_thresholds(out);
}

int vkgv_measure_atomic_rate(vkgv_ctx* c, uint32_t scattered, double* atomics_per_second)
{
    if (!c || !atomics_per_second) return VKGV_ERROR_INVALID_ARGUMENT;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());
    DevBuf keys;
    int rc = dalloc(c, keys, sizeof(unsigned long long) * calib_key_count());
    if (rc) return rc;
    launch_clear_keys(static_cast<unsigned long long*>(keys.p), calib_key_count(), c->stream);
    const int r = measure_atomic_rate(static_cast<unsigned long long*>(keys.p), (uint32_t)c->cus, scattered ? 1u : 0u,
                                      c->stream, atomics_per_second);
    const hipError_t e = hipStreamSynchronize(c->stream);
    dfree(keys);
    if (r != 0 || e != hipSuccess) return fail(c, VKGV_ERROR_HIP, "measure_atomic_rate: a timed launch failed");
    return VKGV_OK;
}

int vkgv_set_tuning(vkgv_ctx* c, const vkgv_tuning* t)
{
    if (!c || !t) return VKGV_ERROR_INVALID_ARGUMENT;
    if (t->small_max_extent < 1 || t->small_max_extent > 64 || t->raster_blocks_per_cu < 1 || t->raster_blocks_per_cu > 16
        || t->tile_item_tris < 1 || t->tile_item_tris > 256 || t->resolve_cache_slots > c->cache_slots)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                    "set_tuning: small_max_extent 1..64, raster_blocks_per_cu 1..16, tile_item_tris 1..256, "
                    "resolve_cache_slots 0.." + std::to_string(c->cache_slots));
    c->tuning.small_max_extent = t->small_max_extent;
    c->tuning.raster_blocks_per_cu = t->raster_blocks_per_cu;
    c->tuning.tile_item_tris = t->tile_item_tris;
    c->tuning.resolve_cache_slots = t->resolve_cache_slots;
    return VKGV_OK;
}

int vkgv_get_tuning(const vkgv_ctx* c, vkgv_tuning* t)
{
    if (!c || !t) return VKGV_ERROR_INVALID_ARGUMENT;
    t->small_max_extent = c->tuning.small_max_extent;
    t->raster_blocks_per_cu = c->tuning.raster_blocks_per_cu;
    t->tile_item_tris = c->tuning.tile_item_tris;
    t->resolve_cache_slots = c->tuning.resolve_cache_slots;
    return VKGV_OK;
}

void vkgv_destroy(vkgv_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& b : c->mesh_bufs) dfree(b);
    dfree(c->prims); dfree(c->materials); dfree(c->cone_base); dfree(c->cones);
    dfree(c->draws); dfree(c->cull); dfree(c->rrec); dfree(c->transforms); dfree(c->srgb);
    dfree(c->srgb_decode); dfree(c->tex_recs); dfree(c->texels); dfree(c->mat_tex); dfree(c->prev_transforms);
    for (DevBuf* b : { &c->an_nodes, &c->an_chan_first, &c->an_channels, &c->an_samplers, &c->an_times, &c->an_values,
                       &c->an_trs, &c->an_chain_first, &c->an_chain })
        dfree(*b);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* vkgv_last_error(const vkgv_ctx* c) { return c ? c->err.c_str() : "null context"; }

int vkgv_sync(vkgv_ctx* c)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

} // extern "C"

// The scene upload of vkgv_upload_scene (mesh arrays in host memory, checked
// on the host) and vkgv_upload_scene_device (mesh arrays in device memory,
// copied device to device and checked by a kernel on the packed copies).
static int upload_scene_impl(vkgv_ctx* c, const vkgv_mesh_desc* meshes, uint32_t mesh_count,
                             const vkgv_primitive_desc* prims, uint32_t prim_count, const vkgv_material* materials,
                             uint32_t material_count, bool device_src)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if ((mesh_count && !meshes) || (prim_count && !prims) || !materials || material_count == 0)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_scene: null array or no materials");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());   // frames of any stream may still read the old buffers
    // ---- validation: everything a kernel dereferences stays in bounds
    for (uint32_t mi = 0; mi < mesh_count; ++mi) {
        const vkgv_mesh_desc& m = meshes[mi];
        if ((m.meshlet_count && !m.meshlets) || (m.vertex_index_count && !m.vertex_indices)
            || (m.triangle_index_bytes && !m.triangle_indices) || (m.vertex_count && !m.vertices))
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_scene: mesh " + std::to_string(mi) + " has a null array");
        if (device_src) continue;         // device arrays: checked on the device after the copy
        for (uint64_t k = 0; k < m.vertex_index_count; ++k)
            if (m.vertex_indices[k] >= m.vertex_count)
                return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_scene: vertex index out of range in mesh " + std::to_string(mi));
        for (uint32_t k = 0; k < m.meshlet_count; ++k) {
            const vkgv_meshlet& ml = m.meshlets[k];
            const uint64_t tbytes = ((uint64_t)ml.triangle_count * 3u + 3u) & ~3ull;
            if (ml.vertex_count == 0 || ml.vertex_count > VKGV_MAX_VERTICES || ml.triangle_count == 0
                || ml.triangle_count > VKGV_MAX_PRIMITIVES || (ml.triangle_offset & 3u)
                || (uint64_t)ml.vertex_offset + ml.vertex_count > m.vertex_index_count
                || (uint64_t)ml.triangle_offset + tbytes > m.triangle_index_bytes)
                return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                            "upload_scene: meshlet " + std::to_string(k) + " of mesh " + std::to_string(mi) + " is malformed");
            for (uint32_t t = 0; t < 3u * ml.triangle_count; ++t)
                if (m.triangle_indices[ml.triangle_offset + t] >= ml.vertex_count)
                    return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                                "upload_scene: triangle index out of range in meshlet " + std::to_string(k) + " of mesh "
                                    + std::to_string(mi));
        }
    }
    for (uint32_t p = 0; p < prim_count; ++p)
        if (prims[p].mesh_index >= mesh_count || prims[p].material_index >= material_count)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_scene: primitive " + std::to_string(p) + " out of range");

    // ---- packing (assets.cpp:375-429 staging upload).  The four buffer
    // kinds of all meshes are packed into one device allocation each (every
    // mesh's part 256-byte aligned); a primitive's four addresses point into
    // them, as the reference's per-primitive buffer addresses do.
    std::vector<uint64_t> off((size_t)mesh_count * 4);
    uint64_t total[4] = { 0, 0, 0, 0 };
    auto part_bytes = [&](const vkgv_mesh_desc& m, int k) -> uint64_t {
        switch (k) {
        case 0: return 4ull * m.vertex_index_count;
        case 1: return m.triangle_index_bytes;
        case 2: return sizeof(vkgv_vertex) * m.vertex_count;
        default: return sizeof(vkgv_meshlet) * (uint64_t)m.meshlet_count;
        }
    };
    auto part_src = [&](const vkgv_mesh_desc& m, int k) -> const void* {
        switch (k) {
        case 0: return m.vertex_indices;
        case 1: return m.triangle_indices;
        case 2: return m.vertices;
        default: return m.meshlets;
        }
    };
    for (uint32_t mi = 0; mi < mesh_count; ++mi)
        for (int k = 0; k < 4; ++k) {
            off[4 * mi + k] = total[k];
            total[k] += (part_bytes(meshes[mi], k) + 255ull) & ~255ull;
        }
    // raster records address the packed buffers with 32-bit offsets (elements
    // for vertex indices, bytes for triangles and vertices)
    for (int k = 0; k < 4; ++k)
        if (total[k] / (k == 0 ? 4u : 1u) >= (1ull << 32))
            return fail(c, VKGV_ERROR_LIMIT, "upload_scene: packed mesh buffer beyond 32-bit offsets");

    // ---- device copies.  From here on the old scene is gone: the context
    // holds no scene (and no draw list, which points into it) until every
    // buffer of the new one is in place, so a failure leaves frames refused
    // with VKGV_ERROR_STATE instead of reading freed memory.
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->scene = false;
    c->draws_set = false;
    c->draw_count = 0;
    c->prim_count = 0;
    c->material_count = 0;
    c->prim_meshlets.clear();
    c->generation++;   // captured batch frames hold the old buffers
    for (auto& b : c->mesh_bufs) dfree(b);
    c->mesh_bufs.assign(4, DevBuf());
    c->mesh_meshlets.assign(mesh_count, 0);
    std::vector<uint32_t> cone_first(mesh_count, 0xFFFFFFFFu);
    std::vector<vkgv_meshlet_cone> all_cones;
    int rc;
    for (int k = 0; k < 4; ++k) {
        if (device_src) {
            // device to device, each mesh's part at its packed offset
            if ((rc = dalloc(c, c->mesh_bufs[k], total[k]))) return rc;
            for (uint32_t mi = 0; mi < mesh_count; ++mi)
                if (part_bytes(meshes[mi], k))
                    HIP_TRY(c, hipMemcpyAsync(static_cast<uint8_t*>(c->mesh_bufs[k].p) + off[4 * mi + k],
                                              part_src(meshes[mi], k), part_bytes(meshes[mi], k), hipMemcpyDefault,
                                              c->stream));
            continue;
        }
        std::vector<uint8_t> staging(total[k] ? total[k] : 1, 0);
        for (uint32_t mi = 0; mi < mesh_count; ++mi)
            if (part_bytes(meshes[mi], k)) std::memcpy(staging.data() + off[4 * mi + k], part_src(meshes[mi], k), part_bytes(meshes[mi], k));
        if ((rc = upload(c, c->mesh_bufs[k], staging.data(), total[k]))) return rc;
        HIP_TRY(c, hipStreamSynchronize(c->stream));   // staging is freed at the end of this iteration
    }
    uint64_t cone_total = 0;
    for (uint32_t mi = 0; mi < mesh_count; ++mi) {
        const vkgv_mesh_desc& m = meshes[mi];
        c->mesh_meshlets[mi] = m.meshlet_count;
        if (m.cones) {
            cone_first[mi] = (uint32_t)cone_total;
            cone_total += m.meshlet_count;
            if (!device_src) all_cones.insert(all_cones.end(), m.cones, m.cones + m.meshlet_count);
        }
    }
    if (cone_total >= 0xFFFFFFFFull) return fail(c, VKGV_ERROR_LIMIT, "upload_scene: more than 2^32 - 1 cones");
    if (device_src) {
        // every index a kernel follows, checked on the packed device copies
        std::vector<MeshCheck> chk(std::max<uint32_t>(mesh_count, 1u));
        uint64_t meshlets = 0;
        for (uint32_t mi = 0; mi < mesh_count; ++mi) {
            const vkgv_mesh_desc& m = meshes[mi];
            MeshCheck& q = chk[mi];
            q.vidx_off = off[4 * mi + 0] / 4u;
            q.vidx_count = m.vertex_index_count;
            q.tri_off = off[4 * mi + 1];
            q.tri_bytes = m.triangle_index_bytes;
            q.vtx_count = m.vertex_count;
            q.ml_off = off[4 * mi + 3];
            q.ml_count = m.meshlet_count;
            q.ml_first = meshlets;
            meshlets += m.meshlet_count;
        }
        DevBuf dchk, dbad;
        if ((rc = upload(c, dchk, chk.data(), sizeof(MeshCheck) * chk.size()))) return rc;
        if ((rc = dalloc(c, dbad, sizeof(unsigned long long)))) { dfree(dchk); return rc; }
        const unsigned long long none = ~0ull;
        unsigned long long bad = none;
        hipError_t e = hipMemcpyAsync(dbad.p, &none, sizeof none, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_check_meshes(static_cast<const MeshCheck*>(dchk.p), mesh_count, meshlets,
                                    static_cast<const uint32_t*>(c->mesh_bufs[0].p),
                                    static_cast<const uint8_t*>(c->mesh_bufs[1].p),
                                    static_cast<const uint8_t*>(c->mesh_bufs[3].p),
                                    static_cast<unsigned long long*>(dbad.p), c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad.p, sizeof bad, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        dfree(dchk);
        dfree(dbad);
        if (e != hipSuccess) return fail(c, VKGV_ERROR_HIP, std::string("upload_scene_device: ") + hipGetErrorString(e));
        if (bad != none)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                        "upload_scene_device: meshlet " + std::to_string(bad & 0xFFFFFFFFull) + " of mesh "
                            + std::to_string(bad >> 32) + " is malformed or indexes out of range");
        if ((rc = dalloc(c, c->cones, sizeof(vkgv_meshlet_cone) * std::max<uint64_t>(cone_total, 1)))) return rc;
        for (uint32_t mi = 0; mi < mesh_count; ++mi)
            if (meshes[mi].cones && meshes[mi].meshlet_count)
                HIP_TRY(c, hipMemcpyAsync(static_cast<vkgv_meshlet_cone*>(c->cones.p) + cone_first[mi], meshes[mi].cones,
                                          sizeof(vkgv_meshlet_cone) * meshes[mi].meshlet_count, hipMemcpyDefault,
                                          c->stream));
    }
    std::vector<vkgv_primitive> dp(prim_count);
    std::vector<uint32_t> cb(std::max<uint32_t>(prim_count, 1u), 0xFFFFFFFFu);
    for (uint32_t p = 0; p < prim_count; ++p) {
        const uint32_t mi = prims[p].mesh_index;
        vkgv_primitive& q = dp[p];
        q.vertex_index_buffer = reinterpret_cast<uint64_t>(c->mesh_bufs[0].p) + off[4 * mi + 0];
        q.primitive_index_buffer = reinterpret_cast<uint64_t>(c->mesh_bufs[1].p) + off[4 * mi + 1];
        q.vertex_buffer = reinterpret_cast<uint64_t>(c->mesh_bufs[2].p) + off[4 * mi + 2];
        q.meshlet_buffer = reinterpret_cast<uint64_t>(c->mesh_bufs[3].p) + off[4 * mi + 3];
        std::memcpy(q.aabb_extents, prims[p].aabb_extents, 12);
        std::memcpy(q.aabb_center, prims[p].aabb_center, 12);
        q.meshlet_count = meshes[mi].meshlet_count;
        q.material_index = prims[p].material_index;
        cb[p] = cone_first[mi];
    }
    if ((rc = upload(c, c->prims, dp.data(), sizeof(vkgv_primitive) * prim_count))) return rc;
    if ((rc = upload(c, c->materials, materials, sizeof(vkgv_material) * material_count))) return rc;
    c->max_albedo_index = -1;
    for (uint32_t m = 0; m < material_count; ++m)
        if (materials[m].albedo_index != 0xFFFFFFFFu)
            c->max_albedo_index = std::max<int64_t>(c->max_albedo_index, materials[m].albedo_index);
    c->host_materials.assign(materials, materials + material_count);
    if ((rc = rebuild_mat_tex(c))) return rc;
    if ((rc = upload(c, c->cone_base, cb.data(), sizeof(uint32_t) * cb.size()))) return rc;
    if (!device_src && (rc = upload(c, c->cones, all_cones.data(), sizeof(vkgv_meshlet_cone) * all_cones.size())))
        return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->prim_meshlets.assign(prim_count, 0);
    for (uint32_t p = 0; p < prim_count; ++p) c->prim_meshlets[p] = meshes[prims[p].mesh_index].meshlet_count;
    c->prim_count = prim_count;
    c->material_count = material_count;
    // a new scene invalidates the draw list (world.cpp:181-182), cleared above
    c->scene = true;
    return VKGV_OK;
}

extern "C" {

int vkgv_upload_scene(vkgv_ctx* c, const vkgv_mesh_desc* meshes, uint32_t mesh_count,
                      const vkgv_primitive_desc* prims, uint32_t prim_count, const vkgv_material* materials,
                      uint32_t material_count)
{
    return upload_scene_impl(c, meshes, mesh_count, prims, prim_count, materials, material_count, false);
}

int vkgv_upload_scene_device(vkgv_ctx* c, const vkgv_mesh_desc* meshes, uint32_t mesh_count,
                             const vkgv_primitive_desc* prims, uint32_t prim_count, const vkgv_material* materials,
                             uint32_t material_count)
{
    return upload_scene_impl(c, meshes, mesh_count, prims, prim_count, materials, material_count, true);
}

int vkgv_upload_textures(vkgv_ctx* c, const vkgv_texture_desc* textures, uint32_t count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (count && !textures) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_textures: null array");
    std::vector<TexRec> recs(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const vkgv_texture_desc& t = textures[i];
        if (!t.rgba8 || t.width == 0 || t.height == 0 || t.width > 16384 || t.height > 16384
            || t.wrap_s > VKGV_WRAP_MIRRORED_REPEAT || t.wrap_t > VKGV_WRAP_MIRRORED_REPEAT)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "upload_textures: texture " + std::to_string(i) + " is malformed");
        recs[i] = TexRec{ total, t.width, t.height, t.wrap_s, t.wrap_t, { 0u, 0u } };
        total += (uint64_t)t.width * t.height;
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());   // frames of any stream may still read the old textures
    c->texture_count = 0;                 // until the new set is complete
    c->generation++;
    // texels stored decoded: toLinear of r, g, b and a / 255, the values the
    // resolve's bilinear filter weights (the same table as the device's)
    int rc = dalloc(c, c->texels, sizeof(float4) * std::max<uint64_t>(total, 1));
    if (rc) return rc;
    {
        float dec[256];
        srgb_decode_table(dec);
        std::vector<float4> lin(std::max<uint64_t>(total, 1));
        for (uint32_t i = 0; i < count; ++i) {
            const uint8_t* src = textures[i].rgba8;
            const uint64_t n = (uint64_t)textures[i].width * textures[i].height;
            for (uint64_t k = 0; k < n; ++k)
                lin[recs[i].offset + k] = make_float4(dec[src[4 * k]], dec[src[4 * k + 1]], dec[src[4 * k + 2]],
                                                      (float)src[4 * k + 3] / 255.0f);
        }
        HIP_TRY(c, hipMemcpyAsync(c->texels.p, lin.data(), sizeof(float4) * lin.size(), hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));   // lin is released here
    }
    if ((rc = upload(c, c->tex_recs, recs.data(), sizeof(TexRec) * count))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->host_textures = recs;
    if ((rc = rebuild_mat_tex(c))) return rc;
    c->texture_count = count;
    c->generation++;   // captured batch frames hold the old texture addresses
    return VKGV_OK;
}

int vkgv_set_draws(vkgv_ctx* c, const vkgv_meshlet_draw* draws, uint32_t count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!c->scene) return fail(c, VKGV_ERROR_STATE, "set_draws: no scene uploaded");
    if (count && !draws) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_draws: null draws");
    if (count > VKGV_MAX_DRAWS) return fail(c, VKGV_ERROR_LIMIT, "set_draws: more than 2^25 draws (25-bit draw index)");
    uint32_t maxt = 0;
    for (uint32_t d = 0; d < count; ++d) {
        const vkgv_meshlet_draw& dr = draws[d];
        if (dr.primitive_index >= c->prim_count || dr.meshlet_index >= c->prim_meshlets[dr.primitive_index])
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_draws: draw " + std::to_string(d) + " out of range");
        maxt = std::max(maxt, dr.transform_index);
    }
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());   // frames of any stream may still read the old buffers
    // no draw list until the new one is complete: a failed allocation leaves
    // frames refused (VKGV_ERROR_STATE), not reading the old list's freed records
    c->draws_set = false;
    c->draw_count = 0;
    c->generation++;
    int rc;
    if ((rc = upload(c, c->draws, draws, sizeof(vkgv_meshlet_draw) * count))) return rc;
    if ((rc = dalloc(c, c->cull, sizeof(CullRec) * std::max<uint32_t>(count, 1u)))) return rc;
    if ((rc = dalloc(c, c->rrec, sizeof(RasterRec) * std::max<uint32_t>(count, 1u)))) return rc;
    c->draw_count = count;
    c->max_transform = maxt;
    c->draws_set = true;
    c->generation++;   // captured batch frames hold the old buffers and grids
    launch_build_cull_records(scene_dev(c), static_cast<CullRec*>(c->cull.p), static_cast<RasterRec*>(c->rrec.p),
                              c->stream);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

namespace {

// Capacity for `count` transforms in the current and previous buffers; when
// the count stays, the current transforms become the previous ones (ordered
// on `st`).  Reallocation synchronizes the device and invalidates captured
// frames.
int reserve_transforms(vkgv_ctx* c, uint32_t count, hipStream_t st)
{
    const size_t bytes = sizeof(float) * 16 * (size_t)count;
    if (c->transforms.bytes < bytes || !c->transforms.p || c->prev_transforms.bytes < bytes || !c->prev_transforms.p) {
        HIP_TRY(c, hipDeviceSynchronize());   // frames of any stream may still read the old buffers
        int rc;
        c->generation++;
        c->prev_count = 0;
        if (c->transforms.bytes < bytes || !c->transforms.p) {
            c->transform_count = 0;           // contents lost, also if the allocation fails
            if ((rc = dalloc(c, c->transforms, bytes))) return rc;
        }
        if ((rc = dalloc(c, c->prev_transforms, std::max(bytes, c->transforms.bytes)))) return rc;
    }
    if (count == c->transform_count && count == c->prev_count && bytes)
        HIP_TRY(c, hipMemcpyAsync(c->prev_transforms.p, c->transforms.p, bytes, hipMemcpyDeviceToDevice, st));
    return VKGV_OK;
}

// After new transforms are in place: a count change leaves no previous frame,
// so the previous transforms are the new ones (zero motion).
int finish_transforms(vkgv_ctx* c, uint32_t count, hipStream_t st)
{
    const size_t bytes = sizeof(float) * 16 * (size_t)count;
    if (!(count == c->transform_count && count == c->prev_count) && bytes)
        HIP_TRY(c, hipMemcpyAsync(c->prev_transforms.p, c->transforms.p, bytes, hipMemcpyDeviceToDevice, st));
    if (count != c->transform_count) c->generation++;   // captured batch frames hold the count
    c->transform_count = count;
    c->prev_count = count;
    return VKGV_OK;
}

} // namespace

int vkgv_set_transforms(vkgv_ctx* c, const float* mat4s, uint32_t count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (count && !mat4s) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_transforms: null array");
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = reserve_transforms(c, count, c->stream);
    if (rc) return rc;
    const size_t bytes = sizeof(float) * 16 * (size_t)count;
    if (bytes) HIP_TRY(c, hipMemcpyAsync(c->transforms.p, mat4s, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = finish_transforms(c, count, c->stream))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

int vkgv_set_transforms_device(vkgv_ctx* c, const float* mat4s, uint32_t count, void* stream)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (count && !mat4s) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_transforms_device: null array");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    int rc = reserve_transforms(c, count, st);
    if (rc) return rc;
    const size_t bytes = sizeof(float) * 16 * (size_t)count;
    if (bytes) HIP_TRY(c, hipMemcpyAsync(c->transforms.p, mat4s, bytes, hipMemcpyDefault, st));
    return finish_transforms(c, count, st);
}

int vkgv_set_hierarchy(vkgv_ctx* c, const vkgv_node* nodes, uint32_t node_count, const uint32_t* children,
                       uint32_t children_count, const uint32_t* roots, uint32_t root_count,
                       const vkgv_anim_sampler* samplers,
                       uint32_t sampler_count, const float* times, uint32_t time_count, const float* values,
                       uint32_t value_count, const vkgv_anim_channel* channels, uint32_t channel_count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if ((node_count && !nodes) || (children_count && !children) || (root_count && !roots) || (sampler_count && !samplers)
        || (time_count && !times) || (value_count && !values) || (channel_count && !channels))
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: null array");
    for (uint32_t n = 0; n < node_count; ++n)
        if ((uint64_t)nodes[n].first_child + nodes[n].child_count > children_count)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                        "set_hierarchy: node " + std::to_string(n) + " lists children past the children array");
    // ---- validation: every index a kernel follows stays in bounds
    for (uint32_t sp = 0; sp < sampler_count; ++sp) {
        const vkgv_anim_sampler& a = samplers[sp];
        if (a.input_count == 0 || (uint64_t)a.input_offset + a.input_count > time_count
            || a.interpolation > VKGV_INTERP_CUBICSPLINE)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: sampler " + std::to_string(sp) + " is malformed");
    }
    std::vector<uint32_t> per_node(node_count + 1, 0);
    for (uint32_t k = 0; k < channel_count; ++k) {
        const vkgv_anim_channel& ch = channels[k];
        if (ch.node >= node_count || ch.sampler >= sampler_count || ch.path > VKGV_PATH_SCALE)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: channel " + std::to_string(k) + " is malformed");
        const vkgv_anim_sampler& a = samplers[ch.sampler];
        const uint64_t comps = ch.path == VKGV_PATH_ROTATION ? 4u : 3u;
        const uint64_t outs = (uint64_t)a.input_count * (a.interpolation == VKGV_INTERP_CUBICSPLINE ? 3u : 1u);
        if ((uint64_t)a.output_offset + outs * comps > value_count)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: channel " + std::to_string(k) + " reads past the values");
        per_node[ch.node + 1]++;
    }
    // channels grouped by node, in array order within a node
    for (uint32_t n = 0; n < node_count; ++n) per_node[n + 1] += per_node[n];
    std::vector<vkgv_anim_channel> grouped(std::max<uint32_t>(channel_count, 1u));
    {
        std::vector<uint32_t> fill(per_node.begin(), per_node.end() - 1);
        for (uint32_t k = 0; k < channel_count; ++k) grouped[fill[channels[k].node]++] = channels[k];
    }
    // depth-first walk of the roots (World::iterateNode order, as vkgv_world_build):
    // every mesh node takes the next transform and records its root-to-node chain
    // glTF node trees form a forest: a node reached twice (listed under two
    // parents, twice under one, or as a root and a child) is rejected, which
    // also rules out cycles and bounds the walk by the node count
    std::vector<uint32_t> chain_first(1, 0), chain;
    std::vector<uint8_t> seen(node_count, 0);
    struct Frame { uint32_t node, depth; };
    std::vector<Frame> stack;
    std::vector<uint32_t> path;
    for (uint32_t r = 0; r < root_count; ++r) {
        if (roots[r] >= node_count) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: root out of range");
        stack.push_back({ roots[r], 0 });
        while (!stack.empty()) {
            const Frame f = stack.back();
            stack.pop_back();
            if (seen[f.node])
                return fail(c, VKGV_ERROR_INVALID_ARGUMENT,
                            "set_hierarchy: node " + std::to_string(f.node) + " is reached twice (not a forest)");
            seen[f.node] = 1;
            if (f.depth >= 64) return fail(c, VKGV_ERROR_LIMIT, "set_hierarchy: deeper than 64 nodes");
            path.resize(f.depth);
            path.push_back(f.node);
            const vkgv_node& n = nodes[f.node];
            if (n.mesh >= 0) {
                chain.insert(chain.end(), path.begin(), path.end());
                chain_first.push_back((uint32_t)chain.size());
            }
            for (uint32_t k = n.child_count; k-- > 0;) {
                const uint32_t child = children[n.first_child + k];
                if (child >= node_count) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "set_hierarchy: child out of range");
                stack.push_back({ child, f.depth + 1 });
            }
        }
    }
    const uint32_t mesh_count = (uint32_t)chain_first.size() - 1;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());   // updates of any stream may still read the old arrays
    c->anim_set = false;
    int rc;
    if ((rc = upload(c, c->an_nodes, nodes, sizeof(vkgv_node) * node_count))) return rc;
    if ((rc = upload(c, c->an_chan_first, per_node.data(), sizeof(uint32_t) * per_node.size()))) return rc;
    if ((rc = upload(c, c->an_channels, grouped.data(), sizeof(vkgv_anim_channel) * channel_count))) return rc;
    if ((rc = upload(c, c->an_samplers, samplers, sizeof(vkgv_anim_sampler) * sampler_count))) return rc;
    if ((rc = upload(c, c->an_times, times, sizeof(float) * time_count))) return rc;
    if ((rc = upload(c, c->an_values, values, sizeof(float) * value_count))) return rc;
    if ((rc = dalloc(c, c->an_trs, sizeof(NodeTrs) * std::max<uint32_t>(node_count, 1u)))) return rc;
    if ((rc = upload(c, c->an_chain_first, chain_first.data(), sizeof(uint32_t) * chain_first.size()))) return rc;
    if ((rc = upload(c, c->an_chain, chain.data(), sizeof(uint32_t) * chain.size()))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    AnimDev& d = c->anim;
    d.nodes = static_cast<const vkgv_node*>(c->an_nodes.p);
    d.node_count = node_count;
    d.chan_first = static_cast<const uint32_t*>(c->an_chan_first.p);
    d.channels = static_cast<const vkgv_anim_channel*>(c->an_channels.p);
    d.samplers = static_cast<const vkgv_anim_sampler*>(c->an_samplers.p);
    d.times = static_cast<const float*>(c->an_times.p);
    d.values = static_cast<const float*>(c->an_values.p);
    d.trs = static_cast<NodeTrs*>(c->an_trs.p);
    d.mesh_count = mesh_count;
    d.chain_first = static_cast<const uint32_t*>(c->an_chain_first.p);
    d.chain_nodes = static_cast<const uint32_t*>(c->an_chain.p);
    // the transform buffer holds the mesh nodes' matrices at time 0 (no previous frame yet)
    if ((rc = reserve_transforms(c, mesh_count, c->stream))) return rc;
    c->prev_count = 0;
    launch_update_transforms(d, 0.0f, static_cast<float*>(c->transforms.p), c->stream);
    HIP_TRY(c, hipGetLastError());
    if ((rc = finish_transforms(c, mesh_count, c->stream))) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->anim_set = true;
    return VKGV_OK;
}

int vkgv_update_transforms(vkgv_ctx* c, float time, void* stream)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!c->anim_set) return fail(c, VKGV_ERROR_STATE, "update_transforms: no hierarchy set");
    if (c->transform_count != c->anim.mesh_count || c->prev_count != c->anim.mesh_count)
        return fail(c, VKGV_ERROR_STATE, "update_transforms: the transforms were replaced since set_hierarchy");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const size_t bytes = sizeof(float) * 16 * (size_t)c->anim.mesh_count;
    if (bytes) HIP_TRY(c, hipMemcpyAsync(c->prev_transforms.p, c->transforms.p, bytes, hipMemcpyDeviceToDevice, st));
    launch_update_transforms(c->anim, time, static_cast<float*>(c->transforms.p), st);
    HIP_TRY(c, hipGetLastError());
    return VKGV_OK;
}

int vkgv_read_transforms(vkgv_ctx* c, float* mat4s, uint32_t count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (count > c->transform_count) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "read_transforms: more than are set");
    if (count && !mat4s) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "read_transforms: null destination");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());   // updates may run on any stream
    if (count) HIP_TRY(c, hipMemcpy(mat4s, c->transforms.p, sizeof(float) * 16 * (size_t)count, hipMemcpyDeviceToHost));
    return VKGV_OK;
}

int vkgv_view_create(vkgv_ctx* c, uint32_t width, uint32_t height, vkgv_view** out)
{
    if (!c || !out) return VKGV_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (width == 0 || height == 0 || width > 16384 || height > 16384 || std::max(width, height) < 2)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "view_create: size must be 1..16384 with max(W,H) >= 2");
    HIP_TRY(c, hipSetDevice(c->device));
    auto* v = new vkgv_view();
    v->ctx = c;
    ViewDev& d = v->dev;
    d.width = width;
    d.height = height;
    // pyramid geometry (application.cpp:472-483)
    const uint32_t m = std::max(width, height);
    uint32_t levels = 0;
    while ((m >> (levels + 1)) != 0) ++levels;
    d.pyr.levels = levels;
    uint32_t off = 0;
    for (uint32_t k = 0; k < 16; ++k) {
        d.pyr.w[k] = d.pyr.h[k] = d.pyr.off[k] = 0;
        if (k < levels) {
            d.pyr.w[k] = std::max(1u, width >> (k + 1));
            d.pyr.h[k] = std::max(1u, height >> (k + 1));
            d.pyr.off[k] = off;
            off += d.pyr.w[k] * d.pyr.h[k];
        }
    }
    const size_t px = (size_t)width * height;
    int rc = VKGV_OK;
    if (!rc) rc = dalloc(c, v->keys, 8 * px);
    if (!rc) rc = dalloc(c, v->color, 4 * px);
    if (!rc) rc = dalloc(c, v->pyramid, 4 * (size_t)std::max(off, 1u));
    if (!rc) rc = dalloc(c, v->keys_b, 8 * px);
    if (!rc) rc = dalloc(c, v->counters, sizeof(Counters));
    if (!rc) rc = dalloc(c, v->counters_b, sizeof(Counters));
    if (!rc) rc = dalloc(c, v->big, sizeof(BigTri) * (size_t)c->big_capacity);
    if (!rc) rc = dalloc(c, v->big_chunk, sizeof(uint32_t) * (size_t)c->big_capacity);
    const uint32_t tiles_x = (width + kBinTile - 1) / kBinTile, tiles_y = (height + kBinTile - 1) / kBinTile;
    const uint32_t ntiles = tiles_x * tiles_y;
    // big-triangle chunk list: 16 chunks (1024 tile slots) per screen tile,
    // 2^15 .. 2^18 chunks (which keeps a wave's chunk sum within 32 bits);
    // per-tile triangle lists: min(1024, 2^24 / tiles), at least 64, unless set
    const uint32_t chunk_cap = std::min<uint32_t>(std::max<uint32_t>(1u << 15, 16u * ntiles), 1u << 18);
    const uint32_t tile_cap = c->tile_capacity ? c->tile_capacity
                                               : std::max<uint32_t>(64u, std::min<uint32_t>(1024u, (1u << 24) / ntiles));
    if ((uint64_t)tile_cap * ntiles > (1ull << 28)) {
        vkgv_view_destroy(v);
        return fail(c, VKGV_ERROR_LIMIT, "view_create: tile_triangle_capacity x screen tiles beyond 2^28");
    }
    if (!rc) rc = dalloc(c, v->clipq, sizeof(ClipTri) * (size_t)c->clip_capacity);
    if (!rc) rc = dalloc(c, v->bin_count, sizeof(uint32_t) * 2u * ((ntiles + 3u) & ~3u));
    if (!rc) rc = dalloc(c, v->bins, sizeof(uint32_t) * (size_t)tile_cap * ntiles);
    if (!rc) rc = dalloc(c, v->setups, sizeof(BigSetup) * (size_t)c->big_capacity);
    if (!rc) rc = dalloc(c, v->tri_cache, sizeof(TriCache) * 128u * (size_t)c->cache_slots);
    if (!rc) rc = dalloc(c, v->slot_albedo, sizeof(float4) * (size_t)c->cache_slots);
    if (!rc) rc = dalloc(c, v->slot_mat, sizeof(uint32_t) * (size_t)c->cache_slots);
    if (!rc) rc = dalloc(c, v->chunk_tri, sizeof(uint32_t) * (size_t)chunk_cap);
    if (!rc) rc = dalloc(c, v->frag_stage, sizeof(FragStage) * (size_t)kFragStageCapacity);
    if (!rc) rc = dalloc(c, v->frag_jobs, sizeof(FragJob) * (size_t)kFragJobCapacity);
    if (rc) { vkgv_view_destroy(v); return rc; }
    d.frag_stage = static_cast<FragStage*>(v->frag_stage.p);
    d.frag_jobs = static_cast<FragJob*>(v->frag_jobs.p);
    d.stage_capacity = kFragStageCapacity;
    d.job_capacity = kFragJobCapacity;
    d.chunk_tri = static_cast<uint32_t*>(v->chunk_tri.p);
    d.setups = static_cast<BigSetup*>(v->setups.p);
    d.tri_cache = static_cast<TriCache*>(v->tri_cache.p);
    d.slot_albedo = static_cast<float4*>(v->slot_albedo.p);
    d.slot_mat = static_cast<uint32_t*>(v->slot_mat.p);
    d.cache_slots = c->cache_slots;
    d.keys = static_cast<unsigned long long*>(v->keys.p);
    d.keys_next = static_cast<unsigned long long*>(v->keys_b.p);
    d.color = static_cast<uint32_t*>(v->color.p);
    d.pyramid = static_cast<float*>(v->pyramid.p);
    d.counters = static_cast<Counters*>(v->counters.p);
    d.counters_next = static_cast<Counters*>(v->counters_b.p);
    d.big = static_cast<BigTri*>(v->big.p);
    d.big_chunk = static_cast<uint32_t*>(v->big_chunk.p);
    d.big_capacity = c->big_capacity;
    d.clipq = static_cast<ClipTri*>(v->clipq.p);
    d.clip_capacity = c->clip_capacity;
    d.tiles_x = tiles_x;
    d.tiles_y = tiles_y;
    d.bin_count = static_cast<uint32_t*>(v->bin_count.p);
    d.bins = static_cast<uint32_t*>(v->bins.p);
    d.tile_capacity = tile_cap;
    d.chunk_capacity = chunk_cap;
    d.hiz_tile_levels = hiz_tile_levels(d.pyr);
    d.hiz_tail_start = hiz_tail_start(d.pyr, d.hiz_tile_levels);
    d.hiz_tiles_x = (d.pyr.w[0] + 31) / 32;
    d.resolve_blocks_x = (width + 63) / 64;
    {
        const std::vector<const ViewDev*> one{ &v->dev };
        if ((rc = build_maps(c, one, v->maps, v->batch, v->grid))) { vkgv_view_destroy(v); return rc; }
    }
    // first frame: pyramid = far (0.0), so the first cull is frustum-only
    // (the reference leaves it undefined, application.cpp:743); empty visbuffer.
    if (hipMemsetAsync(v->pyramid.p, 0, v->pyramid.bytes, c->stream) != hipSuccess
        || hipMemsetAsync(v->counters.p, 0, v->counters.bytes, c->stream) != hipSuccess
        || hipMemsetAsync(v->counters_b.p, 0, v->counters_b.bytes, c->stream) != hipSuccess
        || hipMemsetAsync(v->color.p, 0, v->color.bytes, c->stream) != hipSuccess
        || hipMemsetAsync(v->bin_count.p, 0, v->bin_count.bytes, c->stream) != hipSuccess) {
        vkgv_view_destroy(v);
        return fail(c, VKGV_ERROR_HIP, "view_create: clear failed");
    }
    launch_clear_keys(d.keys, px, c->stream);
    launch_clear_keys(d.keys_next, px, c->stream);
    for (auto& set : v->ev)
        for (auto& e : set)
            if (hipEventCreate(&e) != hipSuccess) { vkgv_view_destroy(v); return fail(c, VKGV_ERROR_HIP, "hipEventCreate"); }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { vkgv_view_destroy(v); return fail(c, VKGV_ERROR_HIP, "sync"); }
    *out = v;
    return VKGV_OK;
}

void vkgv_view_destroy(vkgv_view* v)
{
    if (!v) return;
    vkgv_ctx* c = v->ctx;
    if (c) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        if (v->last_stream) (void)hipStreamSynchronize(v->last_stream);
    }
    for (DevBuf* b : { &v->keys, &v->keys_b, &v->counters_b, &v->color, &v->pyramid, &v->visible0, &v->visible1,
                       &v->history, &v->counters, &v->big,
                       &v->clipq, &v->bin_count, &v->bins, &v->setups, &v->tri_cache,
                       &v->slot_albedo, &v->slot_mat, &v->draw_slot, &v->big_chunk, &v->chunk_tri, &v->maps,
                       &v->aa_mat, &v->mvp_tab, &v->pmvp_tab, &v->motion, &v->redo, &v->frag_stage, &v->frag_jobs })
        dfree(*b);
    for (auto& set : v->ev)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    delete v;
}

int vkgv_render(vkgv_ctx* c, vkgv_view* v, const vkgv_frame_desc* f)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!f) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "render: null frame");
    const uint32_t flags = f->flags;
    if ((rc = check_frame_inputs(c, flags))) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = ensure_lists(c, v))) return rc;
    if ((flags & VKGV_MOTION_VECTORS) && (rc = ensure_motion(c, v))) return rc;
    hipStream_t st = f->stream ? static_cast<hipStream_t>(f->stream) : c->stream;
    if (v->last_stream && v->last_stream != st) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    v->last_stream = st;
    const bool timed = flags & VKGV_TIMING;
    const uint32_t slot = v->ring_next;
    BatchDev b = v->batch;
    b.views = nullptr;
    b.view0 = next_frame_view(v->dev);
    b.cams = nullptr;
    b.cam0.cam = f->camera;
    b.cam0.eye[0] = f->eye[0]; b.cam0.eye[1] = f->eye[1]; b.cam0.eye[2] = f->eye[2];
    frame_launches(c, b, v->grid, flags, st, [&](Ev e) { if (timed) (void)hipEventRecord(v->ev[slot][e], st); });
    v->dev = b.view0;   // the frame's buffers are now the view's current ones
    HIP_TRY(c, hipGetLastError());
    if (timed) {
        v->ring_flags[slot] = flags;
        v->ring_next = (slot + 1) % vkgv_view::kRing;
        v->ring_frames = std::min(v->ring_frames + 1, vkgv_view::kRing);
    }
    v->timed = timed;
    v->last_flags = flags;
    return VKGV_OK;
}

int vkgv_view_build_pyramid(vkgv_ctx* c, vkgv_view* v, void* stream)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    BatchDev b = v->batch;
    b.views = nullptr;
    b.view0 = v->dev;
    b.cams = nullptr;
    launch_hiz(b, v->grid, st);
    HIP_TRY(c, hipGetLastError());
    v->last_stream = st;
    return VKGV_OK;
}

int vkgv_view_reset_history(vkgv_ctx* c, vkgv_view* v)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!v->history.p) return VKGV_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = sync_view(c, v))) return rc;
    HIP_TRY(c, hipMemsetAsync(v->history.p, 0, v->history.bytes, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

int vkgv_view_pyramid_info(const vkgv_view* v, uint32_t* levels, uint32_t* sizes, uint64_t* texels)
{
    if (!v) return VKGV_ERROR_INVALID_ARGUMENT;
    const PyramidDesc& p = v->dev.pyr;
    if (levels) *levels = p.levels;
    uint64_t n = 0;
    for (uint32_t k = 0; k < p.levels; ++k) {
        if (sizes) { sizes[2 * k] = p.w[k]; sizes[2 * k + 1] = p.h[k]; }
        n += (uint64_t)p.w[k] * p.h[k];
    }
    if (texels) *texels = n;
    return VKGV_OK;
}

static int copy_out(vkgv_ctx* c, vkgv_view* v, void* dst, const void* src, size_t bytes)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!dst) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "readback: null destination");
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = sync_view(c, v))) return rc;
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

int vkgv_view_read_visbuffer(vkgv_ctx* c, vkgv_view* v, uint64_t* keys)
{
    if (check_view(c, v)) return check_view(c, v);
    return copy_out(c, v, keys, v->dev.keys, 8ull * v->dev.width * v->dev.height);
}

int vkgv_view_read_ids_depth(vkgv_ctx* c, vkgv_view* v, uint32_t* ids, float* depth)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    const size_t n = (size_t)v->dev.width * v->dev.height;
    std::vector<uint64_t> k(n);
    if ((rc = copy_out(c, v, k.data(), v->dev.keys, 8 * n))) return rc;
    for (size_t i = 0; i < n; ++i) {
        if (ids) ids[i] = 0xFFFFFFFEu - (uint32_t)k[i];
        if (depth) {
            const uint32_t b = ~(uint32_t)(k[i] >> 32);
            std::memcpy(&depth[i], &b, 4);
        }
    }
    return VKGV_OK;
}

static int ensure_aa_mat(vkgv_ctx* c, vkgv_view* v)
{
    if (v->aa_mat.p) return VKGV_OK;
    HIP_TRY(c, hipDeviceSynchronize());   // the allocation must not race work of other streams on the buffer list
    return dalloc(c, v->aa_mat, 4ull * v->dev.width * v->dev.height);
}

static int check_grad(vkgv_ctx* c, vkgv_view* v, const vkgv_camera* cam, const void* a, const void* b)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!cam || !a || !b) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "differentiable frame: null argument");
    if (!c->scene || !c->draws_set) return fail(c, VKGV_ERROR_STATE, "differentiable frame: no scene or draws");
    return check_frame_inputs(c, 0);
}

int vkgv_view_aa_image(vkgv_ctx* c, vkgv_view* v, const vkgv_camera* cam, float* image, void* stream)
{
    int rc = check_grad(c, v, cam, image, image);
    if (rc) return rc;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    if (v->last_stream && v->last_stream != st) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    if ((rc = ensure_aa_mat(c, v))) return rc;
    launch_aa_image(scene_dev(c), v->dev.keys, static_cast<uint32_t*>(v->aa_mat.p), v->dev.width, v->dev.height,
                    cam->view_projection, image, st);
    HIP_TRY(c, hipGetLastError());
    return VKGV_OK;
}

int vkgv_view_aa_image_noalloc(vkgv_ctx* c, vkgv_view* v, const vkgv_camera* cam, float* image, void* stream)
{
    int rc = check_grad(c, v, cam, image, image);
    if (rc) return rc;
    const uint64_t npx = (uint64_t)v->dev.width * v->dev.height;
    if (npx > kNoAllocPixels)
        return fail(c, VKGV_ERROR_LIMIT, "aa_image_noalloc: the view has " + std::to_string(npx)
                                             + " pixels, the scratch holds at most " + std::to_string(kNoAllocPixels));
    if (!c->noalloc_scratch) return fail(c, VKGV_ERROR_STATE, "aa_image_noalloc: no scratch address");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    if (v->last_stream && v->last_stream != st) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    launch_aa_image(scene_dev(c), v->dev.keys, c->noalloc_scratch, v->dev.width, v->dev.height, cam->view_projection,
                    image, st);
    HIP_TRY(c, hipGetLastError());
    return VKGV_OK;
}

int vkgv_view_aa_backward(vkgv_ctx* c, vkgv_view* v, const vkgv_camera* cam, const float* g_image,
                          double* g_transforms, uint32_t transform_count, void* stream)
{
    int rc = check_grad(c, v, cam, g_image, g_transforms);
    if (rc) return rc;
    if (transform_count < c->transform_count)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "aa_backward: g_transforms holds fewer transforms than the scene");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    if (v->last_stream && v->last_stream != st) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    if ((rc = ensure_aa_mat(c, v))) return rc;
    launch_aa_backward(scene_dev(c), v->dev.keys, static_cast<uint32_t*>(v->aa_mat.p), v->dev.width, v->dev.height,
                       cam->view_projection, g_image, g_transforms, st);
    HIP_TRY(c, hipGetLastError());
    return VKGV_OK;
}

int vkgv_view_read_color(vkgv_ctx* c, vkgv_view* v, uint32_t* rgba8)
{
    if (check_view(c, v)) return check_view(c, v);
    return copy_out(c, v, rgba8, v->color.p, 4ull * v->dev.width * v->dev.height);
}

int vkgv_view_read_motion(vkgv_ctx* c, vkgv_view* v, float* motion)
{
    if (check_view(c, v)) return check_view(c, v);
    if (!v->motion.p) return fail(c, VKGV_ERROR_STATE, "read_motion: no frame rendered with VKGV_MOTION_VECTORS");
    return copy_out(c, v, motion, v->motion.p, 8ull * v->dev.width * v->dev.height);
}

int vkgv_view_read_pyramid(vkgv_ctx* c, vkgv_view* v, float* texels)
{
    if (check_view(c, v)) return check_view(c, v);
    uint64_t n = 0;
    vkgv_view_pyramid_info(v, nullptr, nullptr, &n);
    return copy_out(c, v, texels, v->pyramid.p, 4 * n);
}

int vkgv_view_write_pyramid(vkgv_ctx* c, vkgv_view* v, const float* texels)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!texels) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "write_pyramid: null source");
    uint64_t n = 0;
    vkgv_view_pyramid_info(v, nullptr, nullptr, &n);
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = sync_view(c, v))) return rc;
    HIP_TRY(c, hipMemcpyAsync(v->pyramid.p, texels, 4 * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return VKGV_OK;
}

int vkgv_view_write_depth(vkgv_ctx* c, vkgv_view* v, const float* depth)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!depth) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "write_depth: null source");
    HIP_TRY(c, hipSetDevice(c->device));
    if ((rc = sync_view(c, v))) return rc;
    const size_t n = (size_t)v->dev.width * v->dev.height;
    DevBuf tmp;
    if ((rc = upload(c, tmp, depth, 4 * n))) return rc;
    launch_write_depth_keys(v->dev.keys, static_cast<const float*>(tmp.p), n, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    dfree(tmp);
    if (e != hipSuccess) return fail(c, VKGV_ERROR_HIP, std::string("write_depth: ") + hipGetErrorString(e));
    return VKGV_OK;
}

int vkgv_view_read_visible(vkgv_ctx* c, vkgv_view* v, uint32_t pass, uint32_t* out, uint32_t capacity, uint32_t* count)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (pass > 1 || !count) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "read_visible: bad pass or null count");
    Counters h{};
    if ((rc = copy_out(c, v, &h, v->dev.counters, sizeof h))) return rc;
    *count = h.vis_count[pass];
    const uint32_t n = std::min(h.vis_count[pass], capacity);
    if (n && out) {
        const void* src = pass ? v->visible1.p : v->visible0.p;
        if (!src) return fail(c, VKGV_ERROR_STATE, "read_visible: no frame rendered");
        std::vector<RasterRec> recs(n);
        if ((rc = copy_out(c, v, recs.data(), src, sizeof(RasterRec) * n))) return rc;
        for (uint32_t i = 0; i < n; ++i) out[i] = recs[i].draw;
    }
    return VKGV_OK;
}

int vkgv_view_read_stats(vkgv_ctx* c, vkgv_view* v, vkgv_frame_stats* s)
{
    int rc = check_view(c, v);
    if (rc) return rc;
    if (!s) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "read_stats: null");
    Counters h{};
    if ((rc = copy_out(c, v, &h, v->dev.counters, sizeof h))) return rc;
    std::memset(s, 0, sizeof *s);
    s->draws = c->draw_count;
    s->visible_early = h.vis_count[0];
    s->visible_late = h.vis_count[1];
    s->big_triangles = h.big_total_tris;
    s->big_tiles = h.big_total_pairs;
    s->clipped_triangles = h.clipped;
    s->overflow = h.overflow;
    s->max_tile_tris = h.max_tile_tris;
    s->raster_triangles = h.raster_tris;
    s->fragments = h.fragments;
    s->fragment_atomics = h.frag_atomics;
    s->fragment_stores = h.frag_stores;
    // Per-pass milliseconds averaged over the timed frames recorded since
    // the previous read_stats (at most the last 64).
    const uint32_t n = v->ring_frames;
    for (uint32_t f = 0; f < n; ++f) {
        const uint32_t slot = (v->ring_next + vkgv_view::kRing - n + f) % vkgv_view::kRing;
        auto el = [&](Ev a, Ev b) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, v->ev[slot][a], v->ev[slot][b]);
            return ms / (float)n;
        };
        const bool two = v->ring_flags[slot] & VKGV_CULL_TWO_PASS;
        s->ms_clear += el(EV_START, EV_CLEAR);
        s->ms_cull += el(EV_CLEAR, EV_CULL);
        s->ms_raster += el(EV_CULL, EV_RASTER);
        if (two) {
            s->ms_hiz += el(EV_RASTER, EV_HIZ);
            s->ms_cull_late += el(EV_HIZ, EV_CULL_LATE);
            s->ms_raster_late += el(EV_CULL_LATE, EV_RASTER_LATE);
            s->ms_resolve += el(EV_RASTER_LATE, EV_RESOLVE);
        } else {
            s->ms_resolve += el(EV_RASTER, EV_RESOLVE);
        }
        s->ms_pyramid += el(EV_RESOLVE, EV_PYRAMID);
        s->ms_total += el(EV_START, EV_PYRAMID);
    }
    s->timed_frames = n;
    v->ring_frames = 0;
    return VKGV_OK;
}

// ---------------------------------------------------------------------------
// Batches

int vkgv_batch_create(vkgv_ctx* c, vkgv_view* const* views, uint32_t count, vkgv_batch** out)
{
    if (!c || !out) return VKGV_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!views || count == 0 || count > kMaxBatchViews)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "batch_create: 1.." + std::to_string(kMaxBatchViews) + " views");
    for (uint32_t i = 0; i < count; ++i) {
        if (!views[i] || views[i]->ctx != c)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "batch_create: view " + std::to_string(i) + " is not of this context");
        for (uint32_t j = 0; j < i; ++j)
            if (views[j] == views[i]) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "batch_create: a view appears twice");
    }
    HIP_TRY(c, hipSetDevice(c->device));
    auto* b = new vkgv_batch();
    b->ctx = c;
    b->views.assign(views, views + count);
    std::vector<const ViewDev*> devs;
    for (vkgv_view* v : b->views) devs.push_back(&v->dev);
    int rc = build_maps(c, devs, b->maps, b->batch, b->grid);
    if (!rc) rc = dalloc(c, b->cams, sizeof(CameraArg) * count);
    for (uint32_t k = 0; !rc && k < vkgv_batch::kStage; ++k) {
        if (hipHostMalloc(reinterpret_cast<void**>(&b->staging[k]), sizeof(CameraArg) * count, hipHostMallocDefault)
                != hipSuccess
            || hipEventCreateWithFlags(&b->staged[k], hipEventDisableTiming) != hipSuccess)
            rc = fail(c, VKGV_ERROR_HIP, "batch_create: pinned staging");
    }
    for (uint32_t k = 0; !rc && k < vkgv_batch::kRing; ++k)
        if (hipEventCreate(&b->ev0[k]) != hipSuccess || hipEventCreate(&b->ev1[k]) != hipSuccess)
            rc = fail(c, VKGV_ERROR_HIP, "batch_create: events");
    if (rc) { vkgv_batch_destroy(b); return rc; }
    b->batch.cams = static_cast<const CameraArg*>(b->cams.p);
    *out = b;
    return VKGV_OK;
}

void vkgv_batch_destroy(vkgv_batch* b)
{
    if (!b) return;
    vkgv_ctx* c = b->ctx;
    if (c) {
        (void)hipSetDevice(c->device);
        if (b->last_stream) (void)hipStreamSynchronize(b->last_stream);
        (void)hipStreamSynchronize(c->stream);
    }
    for (uint32_t t = 0; t < vkgv_batch::kTables; ++t) {
        if (b->exec[t]) (void)hipGraphExecDestroy(b->exec[t]);
        dfree(b->table[t]);
    }
    for (DevBuf* d : { &b->maps, &b->cams }) dfree(*d);
    for (uint32_t k = 0; k < vkgv_batch::kStage; ++k) {
        if (b->staging[k]) (void)hipHostFree(b->staging[k]);
        if (b->staged[k]) (void)hipEventDestroy(b->staged[k]);
    }
    for (uint32_t k = 0; k < vkgv_batch::kRing; ++k) {
        if (b->ev0[k]) (void)hipEventDestroy(b->ev0[k]);
        if (b->ev1[k]) (void)hipEventDestroy(b->ev1[k]);
    }
    delete b;
}

int vkgv_render_batch(vkgv_ctx* c, vkgv_batch* b, const vkgv_frame_desc* frames, uint32_t count)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!b || b->ctx != c) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "render_batch: batch does not belong to this context");
    if (!frames || count != b->views.size())
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "render_batch: one frame description per view of the batch");
    const uint32_t flags = frames[0].flags;
    for (uint32_t i = 1; i < count; ++i)
        if (frames[i].flags != flags || frames[i].stream != frames[0].stream)
            return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "render_batch: all views of a batch share flags and stream");
    int rc = check_frame_inputs(c, flags);
    if (rc) return rc;
    HIP_TRY(c, hipSetDevice(c->device));
    for (vkgv_view* v : b->views) {
        if ((rc = ensure_lists(c, v))) return rc;
        if ((flags & VKGV_MOTION_VECTORS) && (rc = ensure_motion(c, v))) return rc;
    }
    // the views as this frame uses them, and a device table of them
    std::vector<ViewDev> next;
    std::vector<const void*> keys;
    for (vkgv_view* v : b->views) {
        next.push_back(next_frame_view(v->dev));
        keys.push_back(next.back().keys);
    }
    uint32_t tb = vkgv_batch::kTables;
    for (uint32_t t = 0; t < vkgv_batch::kTables; ++t)
        if (b->table_generation[t] == c->generation && b->table_keys[t] == keys) tb = t;
    if (tb == vkgv_batch::kTables) {
        // an earlier frame may still read the slot: wait for it (only when a
        // table is rebuilt, not in a steady run of frames)
        if (b->last_stream) HIP_TRY(c, hipStreamSynchronize(b->last_stream));
        tb = b->table_next;
        b->table_next = (tb + 1) % vkgv_batch::kTables;
        b->table_generation[tb] = ~0ull;
        if (b->exec[tb]) { (void)hipGraphExecDestroy(b->exec[tb]); b->exec[tb] = nullptr; }
        std::vector<const ViewDev*> devs;
        for (const ViewDev& d : next) devs.push_back(&d);
        BatchDev tmp = b->batch;
        if ((rc = upload_table(c, devs, b->table[tb], tmp))) return rc;
        b->table_keys[tb] = keys;
        b->table_generation[tb] = c->generation;
    }
    BatchDev bd = b->batch;
    bd.views = static_cast<const ViewDev*>(b->table[tb].p);
    hipStream_t st = frames[0].stream ? static_cast<hipStream_t>(frames[0].stream) : c->stream;
    for (vkgv_view* v : b->views)
        if (v->last_stream && v->last_stream != st) HIP_TRY(c, hipStreamSynchronize(v->last_stream));
    // cameras: pinned staging slot -> device (the slot is reused once its copy has run)
    const uint32_t k = b->stage_next;
    if (b->stage_used[k]) HIP_TRY(c, hipEventSynchronize(b->staged[k]));
    for (uint32_t i = 0; i < count; ++i) {
        b->staging[k][i].cam = frames[i].camera;
        std::memcpy(b->staging[k][i].eye, frames[i].eye, sizeof(float) * 3);
    }
    HIP_TRY(c, hipMemcpyAsync(b->cams.p, b->staging[k], sizeof(CameraArg) * count, hipMemcpyHostToDevice, st));
    HIP_TRY(c, hipEventRecord(b->staged[k], st));
    b->stage_used[k] = true;
    b->stage_next = (k + 1) % vkgv_batch::kStage;

    const bool timed = flags & VKGV_TIMING;
    const uint32_t fl = flags & ~(VKGV_TIMING | VKGV_GRAPH);
    const uint32_t slot = b->ring_next;
    if (timed) HIP_TRY(c, hipEventRecord(b->ev0[slot], st));
    if (flags & VKGV_GRAPH) {
        const RasterTuning& et = b->exec_tuning[tb];
        const bool stale = !b->exec[tb] || b->exec_flags[tb] != fl || b->exec_stream[tb] != st
                           || et.small_max_extent != c->tuning.small_max_extent
                           || et.raster_blocks_per_cu != c->tuning.raster_blocks_per_cu
                           || et.tile_item_tris != c->tuning.tile_item_tris
                           || et.resolve_cache_slots != c->tuning.resolve_cache_slots;
        if (stale) {
            if (b->exec[tb]) { (void)hipGraphExecDestroy(b->exec[tb]); b->exec[tb] = nullptr; }
            HIP_TRY(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            frame_launches(c, bd, b->grid, fl, st, [](Ev) {});
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (ce != hipSuccess || !g)
                return fail(c, VKGV_ERROR_HIP, std::string("render_batch: graph capture: ") + hipGetErrorString(ce));
            const hipError_t ie = hipGraphInstantiate(&b->exec[tb], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                b->exec[tb] = nullptr;
                return fail(c, VKGV_ERROR_HIP, std::string("render_batch: graph instantiate: ") + hipGetErrorString(ie));
            }
            b->exec_flags[tb] = fl;
            b->exec_stream[tb] = st;
            b->exec_tuning[tb] = c->tuning;
        }
        HIP_TRY(c, hipGraphLaunch(b->exec[tb], st));
    } else {
        frame_launches(c, bd, b->grid, fl, st, [](Ev) {});
    }
    for (size_t i = 0; i < b->views.size(); ++i) b->views[i]->dev = next[i];   // the frame's buffers are current
    HIP_TRY(c, hipGetLastError());
    if (timed) {
        HIP_TRY(c, hipEventRecord(b->ev1[slot], st));
        b->ring_next = (slot + 1) % vkgv_batch::kRing;
        b->ring_frames = std::min(b->ring_frames + 1, vkgv_batch::kRing);
    }
    b->last_stream = st;
    for (vkgv_view* v : b->views) {
        v->last_stream = st;
        v->last_flags = fl;
    }
    return VKGV_OK;
}

int vkgv_batch_read_stats(vkgv_ctx* c, vkgv_batch* b, vkgv_frame_stats* s)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!b || b->ctx != c || !s) return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "batch_read_stats: bad arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    if (b->last_stream) HIP_TRY(c, hipStreamSynchronize(b->last_stream));
    std::memset(s, 0, sizeof *s);
    for (vkgv_view* v : b->views) {
        Counters h{};
        HIP_TRY(c, hipMemcpy(&h, v->dev.counters, sizeof h, hipMemcpyDeviceToHost));
        s->draws += c->draw_count;
        s->visible_early += h.vis_count[0];
        s->visible_late += h.vis_count[1];
        s->big_triangles += h.big_total_tris;
        s->big_tiles += h.big_total_pairs;
        s->clipped_triangles += h.clipped;
        s->overflow |= h.overflow;
        s->max_tile_tris = std::max(s->max_tile_tris, h.max_tile_tris);
        s->raster_triangles += h.raster_tris;
        s->fragments += h.fragments;
        s->fragment_atomics += h.frag_atomics;
        s->fragment_stores += h.frag_stores;
    }
    const uint32_t n = b->ring_frames;
    for (uint32_t f = 0; f < n; ++f) {
        const uint32_t slot = (b->ring_next + vkgv_batch::kRing - n + f) % vkgv_batch::kRing;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, b->ev0[slot], b->ev1[slot]);
        s->ms_total += ms / (float)n;
    }
    s->timed_frames = n;
    b->ring_frames = 0;
    return VKGV_OK;
}

int vkgv_autotune_batch(vkgv_ctx* c, vkgv_batch* b, const vkgv_frame_desc* frames, uint32_t count,
                        uint32_t frames_per_candidate, vkgv_tuning* best_out, float* best_ms)
{
    if (!c) return VKGV_ERROR_INVALID_ARGUMENT;
    if (!b || b->ctx != c || !frames || count != b->views.size() || !best_out)
        return fail(c, VKGV_ERROR_INVALID_ARGUMENT, "autotune_batch: bad arguments");
    const uint32_t nf = std::max<uint32_t>(frames_per_candidate, 1u);
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t st = frames[0].stream ? static_cast<hipStream_t>(frames[0].stream) : c->stream;
    hipEvent_t e0, e1;
    HIP_TRY(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(c, VKGV_ERROR_HIP, "hipEventCreate"); }
    const RasterTuning before = c->tuning;
    // the candidate values of each knob (the Python autotuner's default set)
    const std::vector<std::vector<uint32_t>> axes = {
        { 4u, 8u, 16u, 32u }, { 2u, 3u }, { 8u, 16u, 32u }, { 0u, c->cache_slots } };
    auto knob = [](RasterTuning& t, int k) -> uint32_t& {
        return k == 0 ? t.small_max_extent : k == 1 ? t.raster_blocks_per_cu : k == 2 ? t.tile_item_tris
                                                                                  : t.resolve_cache_slots;
    };
    std::vector<std::pair<RasterTuning, float>> timed;
    int rc = VKGV_OK;
    // ms per batch frame under t, +inf when a frame had a queue full (the
    // direct path: slower, and a sign that the tuning does not fit the scene)
    auto measure = [&](const RasterTuning& t) -> float {
        for (const auto& e : timed)
            if (std::memcmp(&e.first, &t, sizeof t) == 0) return e.second;
        c->tuning = t;
        float ms = 1e30f;
        bool ok = (rc = vkgv_render_batch(c, b, frames, count)) == VKGV_OK;   // warm-up (and graph capture)
        ok = ok && hipEventRecord(e0, st) == hipSuccess;
        for (uint32_t f = 0; ok && f < nf; ++f) ok = (rc = vkgv_render_batch(c, b, frames, count)) == VKGV_OK;
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
        float el = 0.0f;
        ok = ok && hipEventElapsedTime(&el, e0, e1) == hipSuccess;
        vkgv_frame_stats s;
        ok = ok && (rc = vkgv_batch_read_stats(c, b, &s)) == VKGV_OK;
        if (ok && s.overflow == 0) ms = el / (float)nf;
        if (!ok && rc == VKGV_OK) rc = fail(c, VKGV_ERROR_HIP, "autotune_batch: timing failed");
        timed.push_back({ t, ms });
        return ms;
    };
    RasterTuning cur = before;
    for (int k = 0; k < 4; ++k) {   // start from the current value where it is a candidate
        const auto& ax = axes[k];
        if (std::find(ax.begin(), ax.end(), knob(cur, k)) == ax.end()) knob(cur, k) = ax[0];
    }
    for (int sweep = 0; sweep < 3 && rc == VKGV_OK; ++sweep) {
        bool changed = false;
        for (int k = 0; k < 4 && rc == VKGV_OK; ++k) {
            RasterTuning bt = cur;
            float bm = measure(cur);
            for (uint32_t v : axes[k]) {
                RasterTuning t = cur;
                knob(t, k) = v;
                const float m = measure(t);
                if (m < bm) { bm = m; bt = t; }
            }
            if (std::memcmp(&bt, &cur, sizeof bt) != 0) { cur = bt; changed = true; }
        }
        if (!changed) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != VKGV_OK) { c->tuning = before; return rc; }
    const float m = measure(cur);
    if (m >= 1e30f) {
        c->tuning = before;
        return fail(c, VKGV_ERROR_STATE, "autotune_batch: every candidate filled a raster queue");
    }
    c->tuning = cur;
    best_out->small_max_extent = cur.small_max_extent;
    best_out->raster_blocks_per_cu = cur.raster_blocks_per_cu;
    best_out->tile_item_tris = cur.tile_item_tris;
    best_out->resolve_cache_slots = cur.resolve_cache_slots;
    if (best_ms) *best_ms = m;
    return VKGV_OK;
}

} // extern "C"
