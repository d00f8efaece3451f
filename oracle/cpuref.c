/*
 * cpuref.c - the oracle (test infrastructure only; see cpuref.h).
 *
 * A scalar restatement of the reference's GLSL hot path and of the Vulkan
 * fixed-function rules it relies on.  Compiled with -ffp-contract=off; every
 * float expression below is evaluated left to right as written.
 */
#include "cpuref.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MAT(m, c, r) ((m)[(c) * 4 + (r)])

/* GLSL min/max/clamp (GLSL 4.60 spec 8.3: min(x,y) = y < x ? y : x). */
static inline float gmin(float x, float y) { return y < x ? y : x; }
static inline float gmax(float x, float y) { return x < y ? y : x; }
static inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* ------------------------------------------------------------------------ */
/* Matrices (glm layout: m[c][r]).                                          */

/* C = A * B with each element ((a0*b0 + a1*b1) + a2*b2) + a3*b3. */
void ref_mat4_mul(const float* a, const float* b, float* out)
{
    float t[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            t[c * 4 + r] = ((MAT(a, 0, r) * MAT(b, c, 0) + MAT(a, 1, r) * MAT(b, c, 1))
                            + MAT(a, 2, r) * MAT(b, c, 2)) + MAT(a, 3, r) * MAT(b, c, 3);
    memcpy(out, t, sizeof t);
}

/* out = m * (x, y, z, w) */
static void mat4_vec4(const float* m, const float* v, float* out)
{
    float t[4];
    for (int r = 0; r < 4; ++r)
        t[r] = ((MAT(m, 0, r) * v[0] + MAT(m, 1, r) * v[1]) + MAT(m, 2, r) * v[2]) + MAT(m, 3, r) * v[3];
    memcpy(out, t, sizeof t);
}

/* determinant(mat4), Laplace expansion down column 0; the mesh shader only
 * uses its sign (visbuffer.mesh.glsl:71,94). */
static float det4(const float* m)
{
#define T(ra, rb) (MAT(m, 2, ra) * MAT(m, 3, rb) - MAT(m, 2, rb) * MAT(m, 3, ra))
    const float t23 = T(2, 3), t13 = T(1, 3), t12 = T(1, 2), t03 = T(0, 3), t02 = T(0, 2), t01 = T(0, 1);
#undef T
    const float c0 = (MAT(m, 1, 1) * t23 - MAT(m, 1, 2) * t13) + MAT(m, 1, 3) * t12;
    const float c1 = (MAT(m, 1, 0) * t23 - MAT(m, 1, 2) * t03) + MAT(m, 1, 3) * t02;
    const float c2 = (MAT(m, 1, 0) * t13 - MAT(m, 1, 1) * t03) + MAT(m, 1, 3) * t01;
    const float c3 = (MAT(m, 1, 0) * t12 - MAT(m, 1, 1) * t02) + MAT(m, 1, 2) * t01;
    return ((MAT(m, 0, 0) * c0 - MAT(m, 0, 1) * c1) + MAT(m, 0, 2) * c2) - MAT(m, 0, 3) * c3;
}

/* determinant(mat3(a, b, c)) of three column vectors (visbuffer.mesh.glsl:93). */
static float det3(const float* a, const float* b, const float* c)
{
    return (a[0] * (b[1] * c[2] - c[1] * b[2]) - b[0] * (a[1] * c[2] - c[1] * a[2]))
           + c[0] * (a[1] * b[2] - b[1] * a[2]);
}

/* glm::perspectiveRH_ZO closed form (camera.cpp:177). */
void ref_perspective_rh_zo(float fovy, float aspect, float znear, float zfar, float* out)
{
    const float t = tanf(fovy / 2.0f);
    memset(out, 0, 16 * sizeof(float));
    MAT(out, 0, 0) = 1.0f / (aspect * t);
    MAT(out, 1, 1) = 1.0f / t;
    MAT(out, 2, 2) = zfar / (znear - zfar);
    MAT(out, 2, 3) = -1.0f;
    MAT(out, 3, 2) = -(zfar * znear) / (zfar - znear);
}

/* glm::orthoRH_ZO closed form (camera.cpp:31-33). */
void ref_ortho_rh_zo(float l, float r, float b, float t, float znear, float zfar, float* out)
{
    memset(out, 0, 16 * sizeof(float));
    MAT(out, 0, 0) = 2.0f / (r - l);
    MAT(out, 1, 1) = 2.0f / (t - b);
    MAT(out, 2, 2) = -1.0f / (zfar - znear);
    MAT(out, 3, 0) = -(r + l) / (r - l);
    MAT(out, 3, 1) = -(t + b) / (t - b);
    MAT(out, 3, 2) = -znear / (zfar - znear);
    MAT(out, 3, 3) = 1.0f;
}

static void normalize3(const float* v, float* o)
{
    const float l = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    o[0] = v[0] / l; o[1] = v[1] / l; o[2] = v[2] / l;
}
static void cross3(const float* a, const float* b, float* o)
{
    float t[3] = { a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0] };
    memcpy(o, t, sizeof t);
}
static float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
void ref_mat4_vec4(const float* m, const float* v, float* out) { mat4_vec4(m, v, out); }
void ref_cross3(const float* a, const float* b, float* out) { cross3(a, b, out); }

/* glm::lookAtRH closed form (camera.cpp:170). */
void ref_look_at_rh(const float* eye, const float* center, const float* up, float* out)
{
    float f[3] = { center[0] - eye[0], center[1] - eye[1], center[2] - eye[2] }, s[3], u[3];
    normalize3(f, f);
    cross3(f, up, s);
    normalize3(s, s);
    cross3(s, f, u);
    memset(out, 0, 16 * sizeof(float));
    for (int i = 0; i < 3; ++i) {
        MAT(out, i, 0) = s[i];
        MAT(out, i, 1) = u[i];
        MAT(out, i, 2) = -f[i];
    }
    MAT(out, 3, 0) = -dot3(s, eye);
    MAT(out, 3, 1) = -dot3(u, eye);
    MAT(out, 3, 2) = dot3(f, eye);
    MAT(out, 3, 3) = 1.0f;
}

/* reverseDepth: reverseZ * projection, reverseZ maps z -> w - z (camera.cpp:38-48). */
void ref_reverse_depth(const float* proj, float* out)
{
    float rz[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 1, 1 };
    ref_mat4_mul(rz, proj, out);
}

/* generateCameraFrustum (camera.cpp:70-84): plane i = row3 +/- row k of the
 * view-projection, divided by |xyz|, then w negated. */
void ref_frustum_planes(const float* vp, float* planes)
{
    for (int k = 0; k < 3; ++k) {
        for (int i = 0; i < 4; ++i) {
            planes[(2 * k) * 4 + i]     = MAT(vp, i, 3) + MAT(vp, i, k);
            planes[(2 * k + 1) * 4 + i] = MAT(vp, i, 3) - MAT(vp, i, k);
        }
    }
    for (int p = 0; p < 6; ++p) {
        float* pl = planes + p * 4;
        const float l = sqrtf((pl[0] * pl[0] + pl[1] * pl[1]) + pl[2] * pl[2]);
        pl[0] /= l; pl[1] /= l; pl[2] /= l; pl[3] /= l;
        pl[3] = -pl[3];
    }
}

/* Camera::updateCamera's matrix part (camera.cpp:177-193) for a first frame:
 * P[1][1] *= -1, VP = reverseDepth(P) * view; prev* = current. */
void ref_make_camera(const float* view, const float* proj_no_flip, vkgv_camera* out)
{
    float p[16], rp[16], vp[16];
    memcpy(p, proj_no_flip, sizeof p);
    MAT(p, 1, 1) *= -1.0f;
    ref_reverse_depth(p, rp);
    ref_mat4_mul(rp, view, vp);
    memcpy(out->view_projection, vp, sizeof vp);
    memcpy(out->occlusion_view_projection, vp, sizeof vp);
    memcpy(out->prev_view_projection, vp, sizeof vp);
    memcpy(out->prev_occlusion_view_projection, vp, sizeof vp);
    ref_frustum_planes(vp, &out->frustum[0][0]);
}

/* ------------------------------------------------------------------------ */
/* culling.h.glsl                                                            */

/* isAabbInFrustum (culling.h.glsl:8-19). */
int ref_aabb_in_frustum(const float* center, const float* extents, const float* planes)
{
    for (int i = 0; i < 6; ++i) {
        const float* p = planes + 4 * i;
        const float radius = (extents[0] * fabsf(p[0]) + extents[1] * fabsf(p[1])) + extents[2] * fabsf(p[2]);
        const float distance = ((p[0] * center[0] + p[1] * center[1]) + p[2] * center[2]) - p[3];
        if (-radius > distance)
            return 0;
    }
    return 1;
}

/* World-space AABB: center = (M * vec4(c, 1)).xyz (visbuffer.task.glsl:50),
 * extent = mat3(|M0|,|M1|,|M2|) * e (getWorldSpaceAabbExtent, culling.h.glsl:22-29). */
void ref_world_aabb(const float* m, const float* c, const float* e, float* wc, float* we)
{
    const float v[4] = { c[0], c[1], c[2], 1.0f };
    float t[4];
    mat4_vec4(m, v, t);
    wc[0] = t[0]; wc[1] = t[1]; wc[2] = t[2];
    for (int r = 0; r < 3; ++r)
        we[r] = (fabsf(MAT(m, 0, r)) * e[0] + fabsf(MAT(m, 1, r)) * e[1]) + fabsf(MAT(m, 2, r)) * e[2];
}

/* aabbPositions (culling.h.glsl:32-41). */
static const float k_aabb_corners[8][3] = {
    { 1, -1, -1 }, { 1, 1, -1 }, { -1, 1, -1 }, { -1, -1, -1 },
    { 1, -1, 1 },  { 1, 1, 1 },  { -1, -1, 1 }, { -1, 1, 1 },
};

/* projectAabb (culling.h.glsl:44-56). */
void ref_project_aabb(const float* c, const float* e, const float* vp, float* ss_min, float* ss_max)
{
    float mn[3] = { 1.0f, 1.0f, 1.0f }, mx[3] = { -1.0f, -1.0f, -1.0f };
    for (int i = 0; i < 8; ++i) {
        const float pos[4] = { k_aabb_corners[i][0] * e[0] + c[0], k_aabb_corners[i][1] * e[1] + c[1],
                               k_aabb_corners[i][2] * e[2] + c[2], 1.0f };
        float clip[4];
        mat4_vec4(vp, pos, clip);
        const float nx = gclamp(clip[0] / clip[3], -1.0f, 1.0f);
        const float ny = gclamp(clip[1] / clip[3], -1.0f, 1.0f);
        const float u = nx * 0.5f + 0.5f, v = ny * 0.5f + 0.5f, z = clip[2] / clip[3];
        mn[0] = gmin(mn[0], u); mn[1] = gmin(mn[1], v); mn[2] = gmin(mn[2], z);
        mx[0] = gmax(mx[0], u); mx[1] = gmax(mx[1], v); mx[2] = gmax(mx[2], z);
    }
    memcpy(ss_min, mn, sizeof mn);
    memcpy(ss_max, mx, sizeof mx);
}

/* ------------------------------------------------------------------------ */
/* Hi-Z pyramid                                                              */

/* mipLevels = floor(log2(max(W, H))), base (W>>1, H>>1) (application.cpp:472-483);
 * level k is max(1, W >> (k+1)) x max(1, H >> (k+1)). */
uint32_t ref_pyramid_levels(uint32_t w, uint32_t h, uint32_t* sizes, uint64_t* offsets)
{
    const uint32_t m = w > h ? w : h;
    uint32_t levels = 0;
    while ((m >> (levels + 1)) != 0) ++levels;
    uint64_t off = 0;
    for (uint32_t k = 0; k < levels; ++k) {
        const uint32_t lw = (w >> (k + 1)) ? (w >> (k + 1)) : 1u;
        const uint32_t lh = (h >> (k + 1)) ? (h >> (k + 1)) : 1u;
        if (sizes) { sizes[2 * k] = lw; sizes[2 * k + 1] = lh; }
        if (offsets) offsets[k] = off;
        off += (uint64_t)lw * lh;
    }
    if (offsets) offsets[levels] = off;
    return levels;
}

static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Texels of a linear-filter footprint along one axis for unnormalized u:
 * i0 = floor(u), i1 = i0 + 1 only when its weight frac(u) is non-zero (the
 * MIN reduction takes texels "with non-zero weight"), clamp-to-edge. */
static void axis_footprint(float u, uint32_t size, int32_t* t0, int32_t* t1)
{
    const float fl = floorf(u);
    const float a = u - fl;
    const int32_t i0 = (int32_t)fl;
    *t0 = clampi(i0, 0, (int32_t)size - 1);
    *t1 = a != 0.0f ? clampi(i0 + 1, 0, (int32_t)size - 1) : *t0;
}

/* hiz_reduce.comp.glsl:21-31: texel p of a dst level samples its source at
 * (p + 0.5) / dstSize with the MIN sampler (application.cpp:437-458); the
 * source of level 0 is the depth image, of level k the level k-1. */
void ref_hiz_build(const float* depth, uint32_t w, uint32_t h, float* pyramid)
{
    uint32_t sizes[64];
    uint64_t offs[33];
    const uint32_t levels = ref_pyramid_levels(w, h, sizes, offs);
    const float* src = depth;
    uint32_t sw = w, sh = h;
    for (uint32_t k = 0; k < levels; ++k) {
        const uint32_t dw = sizes[2 * k], dh = sizes[2 * k + 1];
        float* dst = pyramid + offs[k];
#pragma omp parallel for schedule(static) if ((uint64_t)dw * dh >= 65536)
        for (uint32_t y = 0; y < dh; ++y) {
            const float v = (((float)y + 0.5f) / (float)dh) * (float)sh - 0.5f;
            int32_t y0, y1;
            axis_footprint(v, sh, &y0, &y1);
            for (uint32_t x = 0; x < dw; ++x) {
                const float u = (((float)x + 0.5f) / (float)dw) * (float)sw - 0.5f;
                int32_t x0, x1;
                axis_footprint(u, sw, &x0, &x1);
                float d = src[(uint64_t)y0 * sw + x0];
                d = gmin(d, src[(uint64_t)y0 * sw + x1]);
                d = gmin(d, src[(uint64_t)y1 * sw + x0]);
                d = gmin(d, src[(uint64_t)y1 * sw + x1]);
                dst[(uint64_t)y * dw + x] = d;
            }
        }
        src = dst; sw = dw; sh = dh;
    }
}

/* floor(log2(x)) evaluated exactly from the binary exponent (the probe's
 * level, visbuffer.task.glsl:59).  0 -> -inf, NaN/negative -> NaN. */
float ref_floor_log2(float x)
{
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (isinf(x)) return INFINITY;
    int e;
    frexpf(x, &e);            /* x = m * 2^e, m in [0.5, 1) */
    return (float)(e - 1);
}

/* textureLod(pyramid, uv, level) through the MIN sampler: lod clamped to
 * [minLod 0, maxLod 16] (NaN -> 0), NEAREST mip = ceil(lod + 0.5) - 1,
 * clamped to the last level; 2x2 linear footprint at that level. */
float ref_pyramid_probe(const float* pyramid, uint32_t w, uint32_t h, const float* uv, float level)
{
    uint32_t sizes[64];
    uint64_t offs[33];
    const uint32_t levels = ref_pyramid_levels(w, h, sizes, offs);
    float lam = level;
    if (!(lam >= 0.0f)) lam = 0.0f;
    if (lam > 16.0f) lam = 16.0f;
    uint32_t d = (uint32_t)(ceilf(lam + 0.5f) - 1.0f);
    if (d > levels - 1) d = levels - 1;
    const uint32_t lw = sizes[2 * d], lh = sizes[2 * d + 1];
    const float* tex = pyramid + offs[d];
    int32_t x0, x1, y0, y1;
    axis_footprint(uv[0] * (float)lw - 0.5f, lw, &x0, &x1);
    axis_footprint(uv[1] * (float)lh - 0.5f, lh, &y0, &y1);
    float r = tex[(uint64_t)y0 * lw + x0];
    r = gmin(r, tex[(uint64_t)y0 * lw + x1]);
    r = gmin(r, tex[(uint64_t)y1 * lw + x0]);
    r = gmin(r, tex[(uint64_t)y1 * lw + x1]);
    return r;
}

/* visbuffer.task.glsl:56-64 */
int ref_occlusion_visible(const float* wc, const float* we, const float* vp,
                          const float* pyramid, uint32_t w, uint32_t h)
{
    float mn[3], mx[3];
    ref_project_aabb(wc, we, vp, mn, mx);
    const uint32_t pw = (w >> 1) ? (w >> 1) : 1u, ph = (h >> 1) ? (h >> 1) : 1u;  /* textureSize(pyramid, 0) */
    const float width = (mx[0] - mn[0]) * (float)pw;
    const float height = (mx[1] - mn[1]) * (float)ph;
    const float level = ref_floor_log2(gmax(width, height));
    const float uv[2] = { (mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f };
    const float depth = ref_pyramid_probe(pyramid, w, h, uv, level);
    return depth < mx[2];
}

/* Normal-cone test (extension; meshoptimizer's published cone test):
 * culled iff the transform is a proper similarity and
 * dot(normalize(apex' - eye), axis') >= cutoff. */
int ref_cone_culled(const vkgv_meshlet_cone* cone, const float* m, const float* eye)
{
    if (!(cone->cutoff < 1.0f))
        return 0;
    const float* c0 = m; const float* c1 = m + 4; const float* c2 = m + 8;
    const float l0 = dot3(c0, c0), l1 = dot3(c1, c1), l2 = dot3(c2, c2);
    const float tol = l0 * 1e-5f;
    if (!(fabsf(l0 - l1) <= tol && fabsf(l0 - l2) <= tol && fabsf(dot3(c0, c1)) <= tol
          && fabsf(dot3(c0, c2)) <= tol && fabsf(dot3(c1, c2)) <= tol))
        return 0;
    if (!(det4(m) > 0.0f))
        return 0;
    const float a4[4] = { cone->apex[0], cone->apex[1], cone->apex[2], 1.0f };
    const float x4[4] = { cone->axis[0], cone->axis[1], cone->axis[2], 0.0f };
    float apex[4], axis[4];
    mat4_vec4(m, a4, apex);
    mat4_vec4(m, x4, axis);
    normalize3(axis, axis);
    float d[3] = { apex[0] - eye[0], apex[1] - eye[1], apex[2] - eye[2] };
    normalize3(d, d);
    return dot3(d, axis) >= cone->cutoff;
}

static const vkgv_meshlet* draw_meshlet(const ref_scene* s, const vkgv_meshlet_draw* dr,
                                        const vkgv_mesh_desc** mesh_out)
{
    const vkgv_primitive_desc* p = &s->primitives[dr->primitive_index];
    const vkgv_mesh_desc* mesh = &s->meshes[p->mesh_index];
    *mesh_out = mesh;
    return &mesh->meshlets[dr->meshlet_index];
}

static int cull_one(const ref_scene* s, uint32_t d, const vkgv_camera* cam, const float* eye,
                    uint32_t flags)
{
    const vkgv_meshlet_draw* dr = &s->draws[d];
    const vkgv_mesh_desc* mesh;
    const vkgv_meshlet* ml = draw_meshlet(s, dr, &mesh);
    const float* m = s->transforms + 16ull * dr->transform_index;
    float wc[3], we[3];
    ref_world_aabb(m, ml->aabb_center, ml->aabb_extents, wc, we);
    int vis = 1;
    if (flags & VKGV_CULL_FRUSTUM)
        vis = ref_aabb_in_frustum(wc, we, &cam->frustum[0][0]);
    if (vis && (flags & VKGV_CULL_CONE) && mesh->cones && eye)
        vis = !ref_cone_culled(&mesh->cones[dr->meshlet_index], m, eye);
    return vis;
}

/* visbuffer.task.glsl:25-76 per draw (compaction order is irrelevant: the
 * visible list is compared as a set). */
void ref_cull(const ref_scene* s, const vkgv_camera* cam, const float* eye,
              const float* pyramid, uint32_t w, uint32_t h, uint32_t flags, uint8_t* visible)
{
#pragma omp parallel for schedule(static)
    for (uint32_t d = 0; d < s->draw_count; ++d) {
        int vis = cull_one(s, d, cam, eye, flags);
        if (vis && (flags & VKGV_CULL_OCCLUSION_PREV)) {
            const vkgv_meshlet_draw* dr = &s->draws[d];
            const vkgv_mesh_desc* mesh;
            const vkgv_meshlet* ml = draw_meshlet(s, dr, &mesh);
            float wc[3], we[3];
            ref_world_aabb(s->transforms + 16ull * dr->transform_index, ml->aabb_center, ml->aabb_extents, wc, we);
            vis = ref_occlusion_visible(wc, we, cam->prev_occlusion_view_projection, pyramid, w, h);
        }
        visible[d] = (uint8_t)vis;
    }
}

/* ------------------------------------------------------------------------ */
/* Rasterization (Vulkan rules as this implementation fixes them):
 *  - viewport [0,W]x[0,H], depth [0,1] (application.cpp:830-841);
 *  - vertices snapped to 1/256 pixel (8 sub-pixel bits), round-half-even;
 *  - coverage at pixel centres, exact integer edge functions, top-left rule;
 *  - depth = plane through the snapped vertices, tested per pixel against
 *    [0,1] (equivalent to clipping at the far/near planes); triangles with a
 *    vertex at w <= 0 or beyond the 2^21-pixel guard band are clipped in clip
 *    space (near plane z <= w, guard planes) and fanned;
 *  - reversed-Z GREATER_OR_EQUAL against clear 0.0, ties to the later
 *    primitive (application.cpp:326-340, 772-823).                           */

#define GUARD_PX 2097152.0f      /* 2^21 pixels */
#define CLIP_GUARD_PX 1048576.0f /* clip planes at 2^20 pixels */

typedef struct svtx { int64_t x, y; float z; } svtx;

static int project_vtx(const float* c, uint32_t w, uint32_t h, svtx* o)
{
    if (!(c[3] > 0.0f)) return 0;
    const float xd = c[0] / c[3], yd = c[1] / c[3], zd = c[2] / c[3];
    const float hw = (float)w * 0.5f, hh = (float)h * 0.5f;
    const float xf = hw * xd + hw, yf = hh * yd + hh;
    if (!(fabsf(xf) < GUARD_PX && fabsf(yf) < GUARD_PX)) return 0;
    o->x = (int64_t)rintf(xf * 256.0f);
    o->y = (int64_t)rintf(yf * 256.0f);
    o->z = zd;
    return 1;
}

static inline int64_t floor_div256(int64_t v) { return v >= 0 ? v / 256 : -((-v + 255) / 256); }

/* Raster statistics (ref_raster_stats): triangles past the back-face test,
 * and fragments (covered pixel centres with 0 <= z <= 1). */
static uint64_t g_stat_tris, g_stat_frags;

/* *dst = min(*dst, key), atomically: the raster runs its draws on several
 * threads, and a minimum does not depend on the order of the updates. */
static inline void key_min(uint64_t* dst, uint64_t key)
{
    uint64_t cur = __atomic_load_n(dst, __ATOMIC_RELAXED);
    while (key < cur && !__atomic_compare_exchange_n(dst, &cur, key, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

static void fill_triangle(svtx v0, svtx v1, svtx v2, uint32_t id, uint32_t w, uint32_t h, uint64_t* keys)
{
    int64_t area = (v1.x - v0.x) * (v2.y - v0.y) - (v1.y - v0.y) * (v2.x - v0.x);
    if (area == 0) return;
    if (area < 0) { svtx t = v1; v1 = v2; v2 = t; area = -area; }
    int64_t minx = v0.x, maxx = v0.x, miny = v0.y, maxy = v0.y;
    if (v1.x < minx) minx = v1.x; if (v1.x > maxx) maxx = v1.x;
    if (v2.x < minx) minx = v2.x; if (v2.x > maxx) maxx = v2.x;
    if (v1.y < miny) miny = v1.y; if (v1.y > maxy) maxy = v1.y;
    if (v2.y < miny) miny = v2.y; if (v2.y > maxy) maxy = v2.y;
    int64_t px0 = floor_div256(minx - 128 + 255), px1 = floor_div256(maxx - 128);
    int64_t py0 = floor_div256(miny - 128 + 255), py1 = floor_div256(maxy - 128);
    if (px0 < 0) px0 = 0;
    if (py0 < 0) py0 = 0;
    if (px1 > (int64_t)w - 1) px1 = (int64_t)w - 1;
    if (py1 > (int64_t)h - 1) py1 = (int64_t)h - 1;
    if (px0 > px1 || py0 > py1) return;

    const svtx* v[3] = { &v0, &v1, &v2 };
    int64_t ex[3], ey[3], ax[3], ay[3];
    int tl[3];
    for (int k = 0; k < 3; ++k) {
        const svtx* a = v[(k + 1) % 3];
        const svtx* b = v[(k + 2) % 3];
        ex[k] = b->x - a->x; ey[k] = b->y - a->y; ax[k] = a->x; ay[k] = a->y;
        tl[k] = ey[k] < 0 || (ey[k] == 0 && ex[k] > 0);
    }
    const double e1x = (double)(v1.x - v0.x) / 256.0, e1y = (double)(v1.y - v0.y) / 256.0;
    const double e2x = (double)(v2.x - v0.x) / 256.0, e2y = (double)(v2.y - v0.y) / 256.0;
    const double dz1 = (double)v1.z - (double)v0.z, dz2 = (double)v2.z - (double)v0.z;
    const double det = (double)area / 65536.0;
    const float za = (float)((dz1 * e2y - dz2 * e1y) / det);
    const float zb = (float)((dz2 * e1x - dz1 * e2x) / det);

    uint64_t frags = 0;
    for (int64_t py = py0; py <= py1; ++py) {
        const int64_t sy = py * 256 + 128;
        for (int64_t px = px0; px <= px1; ++px) {
            const int64_t sx = px * 256 + 128;
            int inside = 1;
            for (int k = 0; k < 3 && inside; ++k) {
                const int64_t e = ex[k] * (sy - ay[k]) - ey[k] * (sx - ax[k]);
                inside = e > 0 || (e == 0 && tl[k]);
            }
            if (!inside) continue;
            const float dxf = (float)(sx - v0.x) * 0.00390625f;
            const float dyf = (float)(sy - v0.y) * 0.00390625f;
            float z = v0.z + (za * dxf + zb * dyf);
            if (!(z >= 0.0f && z <= 1.0f)) continue;
            z = z + 0.0f;
            const uint64_t key = ((uint64_t)(~f2u(z)) << 32) | (uint64_t)(0xFFFFFFFEu - id);
            key_min(&keys[(uint64_t)py * w + (uint64_t)px], key);
            ++frags;
        }
    }
    if (frags) __atomic_fetch_add(&g_stat_frags, frags, __ATOMIC_RELAXED);
}

void ref_raster_stats(int reset, uint64_t* tris, uint64_t* frags)
{
    if (tris) *tris = __atomic_load_n(&g_stat_tris, __ATOMIC_RELAXED);
    if (frags) *frags = __atomic_load_n(&g_stat_frags, __ATOMIC_RELAXED);
    if (reset) {
        __atomic_store_n(&g_stat_tris, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&g_stat_frags, 0, __ATOMIC_RELAXED);
    }
}

/* Sutherland-Hodgman against a*x + b*y + c*z + d*w >= 0.  A convex polygon
 * gains at most one vertex per plane (3 + 5 <= CLIP_MAX); input that crosses
 * a plane more often (non-finite or degenerate) is dropped (returns 0). */
#define CLIP_MAX 16
static int clip_plane(float (*in)[4], int n, float (*out)[4], float a, float b, float c, float d)
{
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const float* p = in[i];
        const float* q = in[(i + 1) % n];
        const float dp = ((a * p[0] + b * p[1]) + c * p[2]) + d * p[3];
        const float dq = ((a * q[0] + b * q[1]) + c * q[2]) + d * q[3];
        if (dp >= 0.0f) { if (m < CLIP_MAX) memcpy(out[m], p, 16); ++m; }
        if ((dp >= 0.0f) != (dq >= 0.0f)) {
            const float t = dp / (dp - dq);
            if (m < CLIP_MAX)
                for (int k = 0; k < 4; ++k) out[m][k] = p[k] + t * (q[k] - p[k]);
            ++m;
        }
    }
    return m <= CLIP_MAX ? m : 0;
}

void ref_raster_triangle(const float* c0, const float* c1, const float* c2, uint32_t id,
                         uint32_t w, uint32_t h, uint64_t* keys)
{
    svtx s0, s1, s2;
    const int ok0 = project_vtx(c0, w, h, &s0), ok1 = project_vtx(c1, w, h, &s1), ok2 = project_vtx(c2, w, h, &s2);
    if (ok0 && ok1 && ok2) {
        fill_triangle(s0, s1, s2, id, w, h, keys);
        return;
    }
    float bufa[CLIP_MAX][4], bufb[CLIP_MAX][4];
    memcpy(bufa[0], c0, 16); memcpy(bufa[1], c1, 16); memcpy(bufa[2], c2, 16);
    const float gx = CLIP_GUARD_PX / (float)w, gy = CLIP_GUARD_PX / (float)h;
    int n = 3;
    n = clip_plane(bufa, n, bufb, 0.0f, 0.0f, -1.0f, 1.0f);   /* near: w - z >= 0 */
    n = clip_plane(bufb, n, bufa, -1.0f, 0.0f, 0.0f, gx);     /* gx*w - x >= 0 */
    n = clip_plane(bufa, n, bufb, 1.0f, 0.0f, 0.0f, gx);      /* gx*w + x >= 0 */
    n = clip_plane(bufb, n, bufa, 0.0f, -1.0f, 0.0f, gy);     /* gy*w - y >= 0 */
    n = clip_plane(bufa, n, bufb, 0.0f, 1.0f, 0.0f, gy);      /* gy*w + y >= 0 */
    if (n < 3) return;
    svtx sv[CLIP_MAX];
    for (int i = 0; i < n; ++i)
        if (!project_vtx(bufb[i], w, h, &sv[i])) return;
    for (int i = 1; i + 1 < n; ++i)
        fill_triangle(sv[0], sv[i], sv[i + 1], id, w, h, keys);
}

/* Threads the frame functions use (OpenMP); n <= 0 leaves the count as is. */
int ref_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void ref_keys_clear(uint64_t* keys, uint64_t n)
{
#pragma omp parallel for schedule(static) if (n >= 65536)
    for (uint64_t i = 0; i < n; ++i) keys[i] = VKGV_KEY_CLEAR;
}

/* visbuffer.mesh.glsl:30-104 + visbuffer.frag.glsl:36 per listed draw. */
void ref_raster(const ref_scene* s, const float* vp, const uint32_t* draw_list, uint32_t n,
                uint32_t w, uint32_t h, uint32_t flags, uint64_t* keys)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (uint32_t li = 0; li < n; ++li) {
        float clip[VKGV_MAX_VERTICES][4];
        const uint32_t d = draw_list[li];
        const vkgv_meshlet_draw* dr = &s->draws[d];
        const vkgv_primitive_desc* p = &s->primitives[dr->primitive_index];
        const vkgv_mesh_desc* mesh = &s->meshes[p->mesh_index];
        const vkgv_meshlet* ml = &mesh->meshlets[dr->meshlet_index];
        const vkgv_material* mat = &s->materials[p->material_index];
        const float* m = s->transforms + 16ull * dr->transform_index;
        float mvp[16];
        ref_mat4_mul(vp, m, mvp);
        for (uint32_t v = 0; v < ml->vertex_count; ++v) {
            const uint32_t vi = mesh->vertex_indices[ml->vertex_offset + v];
            const float* pos = mesh->vertices[vi].position;
            const float p4[4] = { pos[0], pos[1], pos[2], 1.0f };
            mat4_vec4(mvp, p4, clip[v]);
        }
        const float tdet = det4(m);
        const int cull_faces = !mat->double_sided && !(flags & VKGV_NO_BACKFACE);
        for (uint32_t t = 0; t < ml->triangle_count; ++t) {
            const uint8_t* ti = mesh->triangle_indices + ml->triangle_offset + 3 * t;
            const float* a = clip[ti[0]]; const float* b = clip[ti[1]]; const float* c = clip[ti[2]];
            if (cull_faces) {
                const float va[3] = { a[0], a[1], a[3] }, vb[3] = { b[0], b[1], b[3] }, vc[3] = { c[0], c[1], c[3] };
                const float det = det3(va, vb, vc);
                const int culled = tdet < 0.0f ? det < 0.0f : det > 0.0f;
                if (culled) continue;
            }
            __atomic_fetch_add(&g_stat_tris, 1, __ATOMIC_RELAXED);
            ref_raster_triangle(a, b, c, (d << VKGV_TRIANGLE_BITS) | t, w, h, keys);
        }
    }
}

void ref_keys_depth(const uint64_t* keys, uint64_t n, float* depth)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t b = ~(uint32_t)(keys[i] >> 32);
        memcpy(&depth[i], &b, 4);
    }
}

void ref_keys_ids(const uint64_t* keys, uint64_t n, uint32_t* ids)
{
    for (uint64_t i = 0; i < n; ++i) ids[i] = 0xFFFFFFFEu - (uint32_t)keys[i];
}

/* ------------------------------------------------------------------------ */
/* Resolve                                                                   */

/* fromLinear (srgb.h.glsl:9-15). */
float ref_from_linear(float c)
{
    const int cutoff = c < 0.0031308f;
    const float higher = 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
    const float lower = c * 12.92f;
    return cutoff ? lower : higher;
}

static uint32_t unorm8(float x)
{
    float v = gclamp(x, 0.0f, 1.0f);
    if (!(v >= 0.0f)) v = 0.0f;
    return (uint32_t)(v * 255.0f + 0.5f);
}

/* unorm8(fromLinear(c)) of n values: the resolve's colour encode. */
void ref_srgb_codes(const float* c, uint64_t n, uint8_t* out)
{
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint8_t)unorm8(ref_from_linear(c[i]));
}

/* toLinear (srgb.h.glsl:17-23) of b / 255, as float arithmetic. */
static float to_linear_byte(int b)
{
    const float v = (float)b / 255.0f;
    const int cutoff = v < 0.04045f;
    const float higher = powf((v + 0.055f) / 1.055f, 2.4f);
    const float lower = v / 12.92f;
    return cutoff ? lower : higher;
}

/* IEEE binary16 bits -> float. */
static float half_to_float(uint32_t h)
{
    const uint32_t sgn = (h & 0x8000u) << 16, e = (h >> 10) & 31u, m = h & 1023u;
    uint32_t bits;
    float f;
    if (e == 0u) {
        f = ldexpf((float)m, -24);
        return sgn ? -f : f;
    }
    bits = e == 31u ? (sgn | 0x7F800000u | (m << 13)) : (sgn | ((e - 15u + 127u) << 23) | (m << 13));
    memcpy(&f, &bits, 4);
    return f;
}

/* Texel of integer coordinate k: repeat (glTF 10497), clamp to edge
 * (33071), mirrored repeat (33648). */
static uint32_t wrap_texel(int64_t k, uint32_t n, uint32_t mode)
{
    if (mode == VKGV_WRAP_CLAMP_TO_EDGE) return (uint32_t)(k < 0 ? 0 : (k >= (int64_t)n ? (int64_t)n - 1 : k));
    if (mode == VKGV_WRAP_MIRRORED_REPEAT) {
        /* period 2n: 0..n-1 forwards, then n-1..0 */
        int64_t period = 2 * (int64_t)n, m = ((k % period) + period) % period;
        return (uint32_t)(m < (int64_t)n ? m : period - 1 - m);
    }
    return (uint32_t)((((k % (int64_t)n) + (int64_t)n)) % (int64_t)n);
}

/* Vulkan LINEAR filtering at level 0 along one axis (Vulkan spec,
 * "Texel Filtering"): x = u * n - 1/2, texels floor(x) and floor(x) + 1,
 * weight of the second frac(x).  NaN maps to 0; x is clamped to +-2^30. */
static float texel_pair(float u, uint32_t n, uint32_t mode, uint32_t* i0, uint32_t* i1)
{
    float x = u * (float)n - 0.5f;
    if (x != x) x = 0.0f;
    if (x < -1073741824.0f) x = -1073741824.0f;
    if (x > 1073741824.0f) x = 1073741824.0f;
    const float fl = floorf(x);
    const int64_t k = (int64_t)fl;
    *i0 = wrap_texel(k, n, mode);
    *i1 = wrap_texel(k + 1, n, mode);
    return x - fl;
}

/* The commented-out material sampling of visbuffer.frag.glsl:31-32:
 * c *= toLinear(texture(albedo, transformUv(material, uv))), with
 * transformUv = mat2(cos, -sin, sin, cos) * uv * uvScale + uvOffset
 * (mesh_common.h.glsl:129-135) and the texture's bytes decoded per texel. */
static void sample_albedo(const ref_scene* s, const vkgv_material* mat, const float uv[2], float* c)
{
    const vkgv_texture_desc* tex = &s->textures[mat->albedo_index];
    const float cs = cosf(mat->uv_rotation), sn = sinf(mat->uv_rotation);
    const float ur = cs * uv[0] + sn * uv[1];
    const float vr = (-sn) * uv[0] + cs * uv[1];
    const float tu = ur * mat->uv_scale[0] + mat->uv_offset[0];
    const float tv = vr * mat->uv_scale[1] + mat->uv_offset[1];
    uint32_t x0, x1, y0, y1;
    const float fx = texel_pair(tu, tex->width, tex->wrap_s, &x0, &x1);
    const float fy = texel_pair(tv, tex->height, tex->wrap_t, &y0, &y1);
    const float ax = 1.0f - fx, ay = 1.0f - fy;
    const uint8_t* p00 = tex->rgba8 + 4 * ((size_t)y0 * tex->width + x0);
    const uint8_t* p10 = tex->rgba8 + 4 * ((size_t)y0 * tex->width + x1);
    const uint8_t* p01 = tex->rgba8 + 4 * ((size_t)y1 * tex->width + x0);
    const uint8_t* p11 = tex->rgba8 + 4 * ((size_t)y1 * tex->width + x1);
    for (int j = 0; j < 4; ++j) {
        float t00, t10, t01, t11;
        if (j < 3) {
            t00 = to_linear_byte(p00[j]); t10 = to_linear_byte(p10[j]);
            t01 = to_linear_byte(p01[j]); t11 = to_linear_byte(p11[j]);
        } else {
            t00 = (float)p00[j] / 255.0f; t10 = (float)p10[j] / 255.0f;
            t01 = (float)p01[j] / 255.0f; t11 = (float)p11[j] / 255.0f;
        }
        const float r0 = t00 * ax + t10 * fx;
        const float r1 = t01 * ax + t11 * fx;
        c[j] = c[j] * (r0 * ay + r1 * fy);
    }
}

/* visbuffer_resolve.comp.glsl:17-41; with VKGV_RESOLVE_BARY the colour is
 * albedoFactor * interpolated vertex colour, using perspective-correct
 * barycentrics recomputed from the ids; VKGV_RESOLVE_TEXTURED multiplies in
 * the sampled albedo texture as well. */
void ref_resolve(const ref_scene* s, const float* vp, const uint64_t* keys,
                 uint32_t w, uint32_t h, uint32_t flags, uint32_t* rgba)
{
    const float hw = (float)w * 0.5f, hh = (float)h * 0.5f;
    const float inv_hw = 1.0f / hw, inv_hh = 1.0f / hh;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint32_t py = 0; py < h; ++py) {
        for (uint32_t px = 0; px < w; ++px) {
            const uint64_t i = (uint64_t)py * w + px;
            if (keys[i] == VKGV_KEY_CLEAR) { rgba[i] = 0; continue; }
            const uint32_t id = 0xFFFFFFFEu - (uint32_t)keys[i];
            const uint32_t d = id >> VKGV_TRIANGLE_BITS, t = id & ((1u << VKGV_TRIANGLE_BITS) - 1u);
            const vkgv_meshlet_draw* dr = &s->draws[d];
            const vkgv_primitive_desc* p = &s->primitives[dr->primitive_index];
            const vkgv_material* mat = &s->materials[p->material_index];
            float c[4] = { mat->albedo_factor[0], mat->albedo_factor[1], mat->albedo_factor[2], mat->albedo_factor[3] };
            if (flags & (VKGV_RESOLVE_BARY | VKGV_RESOLVE_TEXTURED)) {
                const vkgv_mesh_desc* mesh = &s->meshes[p->mesh_index];
                const vkgv_meshlet* ml = &mesh->meshlets[dr->meshlet_index];
                const float* m = s->transforms + 16ull * dr->transform_index;
                float mvp[16];
                ref_mat4_mul(vp, m, mvp);
                const uint8_t* ti = mesh->triangle_indices + ml->triangle_offset + 3 * t;
                float hv[3][3];
                float col[3][4], uvs[3][2];
                for (int k = 0; k < 3; ++k) {
                    const vkgv_vertex* vx = &mesh->vertices[mesh->vertex_indices[ml->vertex_offset + ti[k]]];
                    uvs[k][0] = half_to_float(vx->uv[0]);
                    uvs[k][1] = half_to_float(vx->uv[1]);
                    const float p4[4] = { vx->position[0], vx->position[1], vx->position[2], 1.0f };
                    float cl[4];
                    mat4_vec4(mvp, p4, cl);
                    hv[k][0] = cl[0]; hv[k][1] = cl[1]; hv[k][2] = cl[3];
                    for (int j = 0; j < 4; ++j) col[k][j] = (float)vx->color[j] / 255.0f;
                }
                const float xn = (((float)px + 0.5f) - hw) * inv_hw;
                const float yn = (((float)py + 0.5f) - hh) * inv_hh;
                float r[3][3], l[3];
                cross3(hv[1], hv[2], r[0]);
                cross3(hv[2], hv[0], r[1]);
                cross3(hv[0], hv[1], r[2]);
                for (int k = 0; k < 3; ++k) l[k] = (r[k][0] * xn + r[k][1] * yn) + r[k][2];
                const float sum = (l[0] + l[1]) + l[2];
                const float inv = 1.0f / sum;
                for (int k = 0; k < 3; ++k) l[k] = l[k] * inv;
                for (int j = 0; j < 4; ++j)
                    c[j] = c[j] * ((l[0] * col[0][j] + l[1] * col[1][j]) + l[2] * col[2][j]);
                if ((flags & VKGV_RESOLVE_TEXTURED) && mat->albedo_index != 0xFFFFFFFFu) {
                    const float uv[2] = { (l[0] * uvs[0][0] + l[1] * uvs[1][0]) + l[2] * uvs[2][0],
                                          (l[0] * uvs[0][1] + l[1] * uvs[1][1]) + l[2] * uvs[2][1] };
                    sample_albedo(s, mat, uv, c);
                }
            }
            const uint32_t r8 = unorm8(ref_from_linear(c[0]));
            const uint32_t g8 = unorm8(ref_from_linear(c[1]));
            const uint32_t b8 = unorm8(ref_from_linear(c[2]));
            const uint32_t a8 = unorm8(c[3]);
            rgba[i] = r8 | (g8 << 8) | (b8 << 16) | (a8 << 24);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Frames                                                                    */

static uint32_t list_from_flags(const uint8_t* f, uint32_t n, uint32_t* list)
{
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) if (f[i]) list[c++] = i;
    return c;
}

uint32_t ref_frame_reference(const ref_scene* s, const vkgv_camera* cam, const float* eye,
                             uint32_t w, uint32_t h, uint32_t flags,
                             float* pyramid, uint64_t* keys, uint32_t* rgba, uint8_t* visible)
{
    const uint64_t npx = (uint64_t)w * h;
    ref_cull(s, cam, eye, pyramid, w, h, flags, visible);
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (s->draw_count ? s->draw_count : 1));
    const uint32_t n = list_from_flags(visible, s->draw_count, list);
    ref_keys_clear(keys, npx);
    ref_raster(s, cam->view_projection, list, n, w, h, flags, keys);
    if (!(flags & VKGV_SKIP_RESOLVE))
        ref_resolve(s, cam->view_projection, keys, w, h, flags, rgba);
    if (flags & VKGV_BUILD_PYRAMID) {
        float* depth = (float*)malloc(sizeof(float) * npx);
        ref_keys_depth(keys, npx, depth);
        ref_hiz_build(depth, w, h, pyramid);
        free(depth);
    }
    free(list);
    return n;
}

void ref_frame_two_pass(const ref_scene* s, const vkgv_camera* cam, const float* eye,
                        uint32_t w, uint32_t h, uint32_t flags,
                        uint8_t* history, float* pyramid, uint64_t* keys, uint32_t* rgba,
                        uint8_t* visible_early, uint8_t* visible_late)
{
    const uint64_t npx = (uint64_t)w * h;
    const uint32_t nd = s->draw_count;
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (nd ? nd : 1));
    uint8_t* pass = (uint8_t*)malloc(nd ? nd : 1);
#pragma omp parallel for schedule(static)
    for (uint32_t d = 0; d < nd; ++d) {
        pass[d] = (uint8_t)cull_one(s, d, cam, eye, flags);
        visible_early[d] = (uint8_t)(pass[d] && history[d]);
    }
    ref_keys_clear(keys, npx);
    uint32_t n = list_from_flags(visible_early, nd, list);
    ref_raster(s, cam->view_projection, list, n, w, h, flags, keys);

    float* depth = (float*)malloc(sizeof(float) * npx);
    ref_keys_depth(keys, npx, depth);
    ref_hiz_build(depth, w, h, pyramid);
    free(depth);

#pragma omp parallel for schedule(static)
    for (uint32_t d = 0; d < nd; ++d) {
        int vis = 0;
        if (pass[d]) {
            const vkgv_meshlet_draw* dr = &s->draws[d];
            const vkgv_mesh_desc* mesh;
            const vkgv_meshlet* ml = draw_meshlet(s, dr, &mesh);
            float wc[3], we[3];
            ref_world_aabb(s->transforms + 16ull * dr->transform_index, ml->aabb_center, ml->aabb_extents, wc, we);
            vis = ref_occlusion_visible(wc, we, cam->occlusion_view_projection, pyramid, w, h);
        }
        visible_late[d] = (uint8_t)(vis && !visible_early[d]);
        history[d] = (uint8_t)vis;
    }
    n = list_from_flags(visible_late, nd, list);
    ref_raster(s, cam->view_projection, list, n, w, h, flags, keys);
    if (!(flags & VKGV_SKIP_RESOLVE))
        ref_resolve(s, cam->view_projection, keys, w, h, flags, rgba);
    free(pass);
    free(list);
}
