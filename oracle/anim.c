/* Oracle (test infrastructure only) for the per-frame transform update and
 * the motion vectors: a scalar restatement of
 *   - AnimationSampler::sample (assets.hpp:52-110): fmod loop, lower_bound,
 *     STEP / LINEAR (slerp for rotations) / CUBICSPLINE (glTF appendix C);
 *   - World::iterateNode (world.cpp:187-228): recursive depth-first walk,
 *     every animation channel naming the node applied in order, then
 *     matrix = parent * T * R * S; mesh nodes take transforms in visit order;
 *   - the fragment shader's motion vectors (visbuffer.frag.glsl:38).
 * fastgltf (the reference's math and slerp) is not vendored with the
 * reference; the slerp here is the textbook shortest-arc formula, with acos
 * and sin evaluated by fixed double-precision series (the same series the
 * device kernels use), and the cubic basis in the glTF specification's form.
 * The functions are pinned by closed forms in tests/test_oracle_anim.py. */
#include "cpuref.h"

#include <math.h>
#include <string.h>

static double series_asin(double y)
{
    const double y2 = y * y;
    double acc = 1.0;
    for (int n = 29; n >= 0; --n)
        acc = 1.0 + (y2 * acc) * ((double)((2 * n + 1) * (2 * n + 1)) / (double)((2 * n + 2) * (2 * n + 3)));
    return y * acc;
}

static double series_acos(double x)
{
    if (x <= 0.5) return 1.5707963267948966 - series_asin(x);
    return 2.0 * series_asin(sqrt((1.0 - x) * 0.5));
}

static double series_sin(double a)
{
    const double a2 = a * a;
    double acc = 1.0;
    for (int n = 14; n >= 0; --n) acc = 1.0 - (a2 * acc) / (double)((2 * n + 2) * (2 * n + 3));
    return a * acc;
}

double ref_series_acos(double x) { return series_acos(x); }
double ref_series_sin(double a) { return series_sin(a); }

static void out_k(const float* v, uint32_t k, int n, float* o)
{
    for (int i = 0; i < n; ++i) o[i] = v[(size_t)k * n + i];
}

/* AnimationSampler::sample; n = 3 (translation, scale) or 4 (rotation). */
void ref_anim_sample(const vkgv_anim_sampler* sp, const float* times, const float* values, int n, float time,
                     float* out)
{
    const float* in = times + sp->input_offset;
    const float* val = values + sp->output_offset;
    const uint32_t cnt = sp->input_count;
    const int cubic = sp->interpolation == VKGV_INTERP_CUBICSPLINE;
    time = fmodf(time, in[cnt - 1]);
    /* std::lower_bound (libstdc++'s bisection: first, len) */
    uint32_t i = 0, len = cnt;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (in[i + half] < time) { i = i + half + 1; len = len - half - 1; }
        else len = half;
    }
    if (i == 0) { out_k(val, cubic ? 1u : 0u, n, out); return; }
    if (i == cnt) { out_k(val, cubic ? 3u * cnt - 2u : cnt - 1u, n, out); return; }
    const float t = (time - in[i - 1]) / (in[i] - in[i - 1]);
    if (sp->interpolation == VKGV_INTERP_STEP) { out_k(val, i - 1, n, out); return; }
    if (!cubic) {
        float a[4] = { 0, 0, 0, 0 }, b[4] = { 0, 0, 0, 0 };
        out_k(val, i - 1, n, a);
        out_k(val, i, n, b);
        if (n == 3) {
            for (int k = 0; k < 3; ++k) out[k] = a[k] * (1.0f - t) + b[k] * t;
            return;
        }
        float c = a[0] * b[0];
        for (int k = 1; k < 4; ++k) c = c + a[k] * b[k];
        if (c < 0.0f) {                      /* shortest arc */
            for (int k = 0; k < 4; ++k) b[k] = -b[k];
            c = -c;
        }
        if (c > 1.0f - 1.1920928955078125e-7f) {
            for (int k = 0; k < 4; ++k) out[k] = a[k] * (1.0f - t) + b[k] * t;
            return;
        }
        const double ang = series_acos((double)c);
        const double sa = series_sin(ang);
        const double w0 = series_sin((1.0 - (double)t) * ang) / sa;
        const double w1 = series_sin((double)t * ang) / sa;
        for (int k = 0; k < 4; ++k) out[k] = (float)(w0 * (double)a[k] + w1 * (double)b[k]);
        return;
    }
    const float t2 = t * t, t3 = t2 * t;
    const float dt = in[i] - in[i - 1];
    float v0[4], b0[4], v1[4], a1[4];
    out_k(val, 3u * (i - 1) + 1u, n, v0);
    out_k(val, 3u * (i - 1) + 2u, n, b0);
    out_k(val, 3u * i + 1u, n, v1);
    out_k(val, 3u * i + 0u, n, a1);
    const float h00 = (2.0f * t3 - 3.0f * t2) + 1.0f;
    const float h10 = (t3 - 2.0f * t2) + t;
    const float h01 = -2.0f * t3 + 3.0f * t2;
    const float h11 = t3 - t2;
    for (int k = 0; k < n; ++k) out[k] = ((v0[k] * h00 + (b0[k] * h10) * dt) + v1[k] * h01) + (a1[k] * h11) * dt;
    if (n == 4) {
        float l2 = out[0] * out[0];
        for (int k = 1; k < 4; ++k) l2 = l2 + out[k] * out[k];
        const float l = sqrtf(l2);
        for (int k = 0; k < 4; ++k) out[k] = out[k] / l;
    }
}

typedef struct {
    const vkgv_node* nodes;
    uint32_t node_count;
    const uint32_t* children;
    const vkgv_anim_sampler* samplers;
    const float* times;
    const float* values;
    const vkgv_anim_channel* channels;
    uint32_t channel_count;
    float time;
    float* out;
    uint32_t count;
    int depth;
    int error;
} walk_t;

static void trs_mul(const float* parent, const float* t, const float* q, const float* s, float* out)
{
    float T[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, t[0], t[1], t[2], 1 };
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    float R[16] = { 1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y), 0.0f,
                    2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x), 0.0f,
                    2.0f * (x * z + w * y), 2.0f * (y * z - w * x), 1.0f - 2.0f * (x * x + y * y), 0.0f,
                    0.0f, 0.0f, 0.0f, 1.0f };
    float S[16] = { s[0], 0, 0, 0, 0, s[1], 0, 0, 0, 0, s[2], 0, 0, 0, 0, 1 };
    float a[16], b[16];
    ref_mat4_mul(parent, T, a);
    ref_mat4_mul(a, R, b);
    ref_mat4_mul(b, S, out);
}

static void visit(walk_t* w, uint32_t node, const float* parent)
{
    if (w->error) return;
    if (node >= w->node_count || w->depth >= 64) { w->error = 1; return; }
    const vkgv_node* n = &w->nodes[node];
    float t[3], q[4], s[3];
    memcpy(t, n->translation, sizeof t);
    memcpy(q, n->rotation, sizeof q);
    memcpy(s, n->scale, sizeof s);
    for (uint32_t k = 0; k < w->channel_count; ++k) {        /* world.cpp:194-216 */
        const vkgv_anim_channel* ch = &w->channels[k];
        if (ch->node != node) continue;
        const vkgv_anim_sampler* sp = &w->samplers[ch->sampler];
        if (ch->path == VKGV_PATH_TRANSLATION) ref_anim_sample(sp, w->times, w->values, 3, w->time, t);
        else if (ch->path == VKGV_PATH_ROTATION) ref_anim_sample(sp, w->times, w->values, 4, w->time, q);
        else ref_anim_sample(sp, w->times, w->values, 3, w->time, s);
    }
    float m[16];
    trs_mul(parent, t, q, s, m);
    if (n->mesh >= 0) {
        if (w->out) memcpy(w->out + 16ull * w->count, m, sizeof m);
        w->count++;
    }
    w->depth++;
    for (uint32_t k = 0; k < n->child_count; ++k) visit(w, w->children[n->first_child + k], m);
    w->depth--;
}

int ref_update_transforms(const vkgv_node* nodes, uint32_t node_count, const uint32_t* children,
                          const uint32_t* roots, uint32_t root_count, const vkgv_anim_sampler* samplers,
                          const float* times, const float* values, const vkgv_anim_channel* channels,
                          uint32_t channel_count, float time, float* out, uint32_t* count)
{
    walk_t w = { nodes, node_count, children, samplers, times, values, channels, channel_count, time, out, 0, 0, 0 };
    const float identity[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    for (uint32_t r = 0; r < root_count; ++r) visit(&w, roots[r], identity);
    *count = w.count;
    return w.error ? -1 : 0;
}

/* Motion vectors (visbuffer.frag.glsl:38) of a frame's keys: per covered
 * pixel, the perspective-correct barycentrics of the resolve (current VP * M),
 * the vertices through prevVP * previous M, interpolated, against the pixel
 * centre's NDC; (0, 0) elsewhere. */
void ref_motion_vectors(const ref_scene* s, const vkgv_camera* cam, const float* prev_transforms,
                        const uint64_t* keys, uint32_t w, uint32_t h, float* motion)
{
    const float hw = (float)w * 0.5f, hh = (float)h * 0.5f;
    const float inv_hw = 1.0f / hw, inv_hh = 1.0f / hh;
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            const uint64_t i = (uint64_t)py * w + px;
            motion[2 * i] = motion[2 * i + 1] = 0.0f;
            if (keys[i] == VKGV_KEY_CLEAR) continue;
            const uint32_t id = 0xFFFFFFFEu - (uint32_t)keys[i];
            const uint32_t d = id >> VKGV_TRIANGLE_BITS, t = id & ((1u << VKGV_TRIANGLE_BITS) - 1u);
            if (d >= s->draw_count) continue;
            const vkgv_meshlet_draw* dr = &s->draws[d];
            const vkgv_primitive_desc* p = &s->primitives[dr->primitive_index];
            const vkgv_mesh_desc* mesh = &s->meshes[p->mesh_index];
            const vkgv_meshlet* ml = &mesh->meshlets[dr->meshlet_index];
            float mvp[16], pmvp[16];
            ref_mat4_mul(cam->view_projection, s->transforms + 16ull * dr->transform_index, mvp);
            ref_mat4_mul(cam->prev_view_projection, prev_transforms + 16ull * dr->transform_index, pmvp);
            const uint8_t* ti = mesh->triangle_indices + ml->triangle_offset + 3 * t;
            float hv[3][3], pq[3][4];
            for (int k = 0; k < 3; ++k) {
                const vkgv_vertex* vx = &mesh->vertices[mesh->vertex_indices[ml->vertex_offset + ti[k]]];
                const float p4[4] = { vx->position[0], vx->position[1], vx->position[2], 1.0f };
                float cl[4];
                ref_mat4_vec4(mvp, p4, cl);
                hv[k][0] = cl[0]; hv[k][1] = cl[1]; hv[k][2] = cl[3];
                ref_mat4_vec4(pmvp, p4, pq[k]);
            }
            const float xn = (((float)px + 0.5f) - hw) * inv_hw;
            const float yn = (((float)py + 0.5f) - hh) * inv_hh;
            float r[3][3], l[3];
            ref_cross3(hv[1], hv[2], r[0]);
            ref_cross3(hv[2], hv[0], r[1]);
            ref_cross3(hv[0], hv[1], r[2]);
            for (int k = 0; k < 3; ++k) l[k] = (r[k][0] * xn + r[k][1] * yn) + r[k][2];
            const float inv = 1.0f / ((l[0] + l[1]) + l[2]);
            for (int k = 0; k < 3; ++k) l[k] = l[k] * inv;
            const float qx = (l[0] * pq[0][0] + l[1] * pq[1][0]) + l[2] * pq[2][0];
            const float qy = (l[0] * pq[0][1] + l[1] * pq[1][1]) + l[2] * pq[2][1];
            const float qw = (l[0] * pq[0][3] + l[1] * pq[1][3]) + l[2] * pq[2][3];
            motion[2 * i] = ((qx / qw) * 0.5f - 0.5f) - (xn * 0.5f - 0.5f);
            motion[2 * i + 1] = ((qy / qw) * 0.5f - 0.5f) - (yn * 0.5f - 0.5f);
        }
}
