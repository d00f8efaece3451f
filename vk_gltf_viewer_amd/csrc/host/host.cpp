// libvkgv_host.so: host-side producers of the frame's inputs (no GPU code).
//
//  - camera math: Camera::updateCamera (camera.cpp:108-194)
//  - World: node walk, draw list, transforms (world.cpp:187-345)
//  - meshlet builder at the reference's limits (assets.cpp:322-373)
//  - seeded synthetic scenes for the benchmark configurations
//
// Compiled with -ffp-contract=off so that the camera matrices are the same
// bits as any other restatement of the same closed forms.
#include "vkgv_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_error;

// ---------------------------------------------------------------------------
// Small vector / matrix helpers (glm layout: m[c*4 + r]).
struct V3 { float x, y, z; };
inline V3 operator+(V3 a, V3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
inline V3 operator-(V3 a, V3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline V3 operator*(V3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { float l = length(a); return { a.x / l, a.y / l, a.z / l }; }
inline V3 vmin(V3 a, V3 b) { return { b.x < a.x ? b.x : a.x, b.y < a.y ? b.y : a.y, b.z < a.z ? b.z : a.z }; }
inline V3 vmax(V3 a, V3 b) { return { a.x < b.x ? b.x : a.x, a.y < b.y ? b.y : a.y, a.z < b.z ? b.z : a.z }; }

using M4 = std::array<float, 16>;
inline float& at(M4& m, int c, int r) { return m[c * 4 + r]; }
inline float at(const M4& m, int c, int r) { return m[c * 4 + r]; }

M4 identity() { M4 m{}; m[0] = m[5] = m[10] = m[15] = 1.0f; return m; }

M4 mul(const M4& a, const M4& b)
{
    M4 t{};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            t[c * 4 + r] = ((at(a, 0, r) * at(b, c, 0) + at(a, 1, r) * at(b, c, 1)) + at(a, 2, r) * at(b, c, 2))
                           + at(a, 3, r) * at(b, c, 3);
    return t;
}

M4 perspective_rh_zo(float fovy, float aspect, float n, float f)
{
    const float t = std::tan(fovy / 2.0f);
    M4 m{};
    at(m, 0, 0) = 1.0f / (aspect * t);
    at(m, 1, 1) = 1.0f / t;
    at(m, 2, 2) = f / (n - f);
    at(m, 2, 3) = -1.0f;
    at(m, 3, 2) = -(f * n) / (f - n);
    return m;
}

M4 ortho_rh_zo(float l, float r, float b, float t, float n, float f)
{
    M4 m{};
    at(m, 0, 0) = 2.0f / (r - l);
    at(m, 1, 1) = 2.0f / (t - b);
    at(m, 2, 2) = -1.0f / (f - n);
    at(m, 3, 0) = -(r + l) / (r - l);
    at(m, 3, 1) = -(t + b) / (t - b);
    at(m, 3, 2) = -n / (f - n);
    at(m, 3, 3) = 1.0f;
    return m;
}

M4 look_at_rh(V3 eye, V3 center, V3 up)
{
    const V3 f = normalize(center - eye);
    const V3 s = normalize(cross(f, up));
    const V3 u = cross(s, f);
    M4 m{};
    at(m, 0, 0) = s.x; at(m, 1, 0) = s.y; at(m, 2, 0) = s.z;
    at(m, 0, 1) = u.x; at(m, 1, 1) = u.y; at(m, 2, 1) = u.z;
    at(m, 0, 2) = -f.x; at(m, 1, 2) = -f.y; at(m, 2, 2) = -f.z;
    at(m, 3, 0) = -dot(s, eye);
    at(m, 3, 1) = -dot(u, eye);
    at(m, 3, 2) = dot(f, eye);
    at(m, 3, 3) = 1.0f;
    return m;
}

// reverseZ maps clip z to w - z (camera.cpp:38-48).
M4 reverse_depth(const M4& p)
{
    const M4 rz = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, 0, 1, 1 };
    return mul(rz, p);
}

// Gribb-Hartmann planes, normalized by |xyz|, w negated (camera.cpp:70-84).
void frustum_planes(const M4& vp, float planes[6][4])
{
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 4; ++i) {
            planes[2 * k][i] = at(vp, i, 3) + at(vp, i, k);
            planes[2 * k + 1][i] = at(vp, i, 3) - at(vp, i, k);
        }
    for (int p = 0; p < 6; ++p) {
        float* pl = planes[p];
        const float l = std::sqrt((pl[0] * pl[0] + pl[1] * pl[1]) + pl[2] * pl[2]);
        pl[0] /= l; pl[1] /= l; pl[2] /= l; pl[3] /= l;
        pl[3] = -pl[3];
    }
}

M4 view_projection(const vkgv_view_desc& v)
{
    const V3 eye{ v.eye[0], v.eye[1], v.eye[2] }, tgt{ v.target[0], v.target[1], v.target[2] },
        up{ v.up[0], v.up[1], v.up[2] };
    const M4 view = look_at_rh(eye, tgt, up);
    M4 proj;
    if (v.orthographic)
        proj = ortho_rh_zo(-v.ortho_half_width, v.ortho_half_width, -v.ortho_half_height, v.ortho_half_height,
                           v.znear, v.zfar);
    else
        proj = perspective_rh_zo(v.fovy, static_cast<float>(v.width) / static_cast<float>(v.height), v.znear, v.zfar);
    at(proj, 1, 1) *= -1.0f;
    return mul(reverse_depth(proj), view);
}

// Column-major TRS: parent * T * R * S (world.cpp:221).
M4 trs(const M4& parent, const float* t, const float* q, const float* s)
{
    M4 T = identity();
    at(T, 3, 0) = t[0]; at(T, 3, 1) = t[1]; at(T, 3, 2) = t[2];
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    M4 R = identity();
    at(R, 0, 0) = 1.0f - 2.0f * (y * y + z * z);
    at(R, 0, 1) = 2.0f * (x * y + w * z);
    at(R, 0, 2) = 2.0f * (x * z - w * y);
    at(R, 1, 0) = 2.0f * (x * y - w * z);
    at(R, 1, 1) = 1.0f - 2.0f * (x * x + z * z);
    at(R, 1, 2) = 2.0f * (y * z + w * x);
    at(R, 2, 0) = 2.0f * (x * z + w * y);
    at(R, 2, 1) = 2.0f * (y * z - w * x);
    at(R, 2, 2) = 1.0f - 2.0f * (x * x + y * y);
    M4 S = identity();
    at(S, 0, 0) = s[0]; at(S, 1, 1) = s[1]; at(S, 2, 2) = s[2];
    return mul(mul(mul(parent, T), R), S);
}

// ---------------------------------------------------------------------------
// Seeded RNG: splitmix64 seeds a PCG32 (XSH-RR) stream.
struct Rng {
    uint64_t state, inc;
    static uint64_t splitmix(uint64_t& x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t s = seed;
        state = splitmix(s);
        inc = splitmix(s) | 1u;
    }
    uint32_t next()
    {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    float uniform() { return static_cast<float>(next() >> 8) * (1.0f / 16777216.0f); }   // [0,1)
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    double normal()
    {
        double u1 = (static_cast<double>(next()) + 1.0) / 4294967297.0;
        double u2 = static_cast<double>(next()) / 4294967296.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// ---------------------------------------------------------------------------
// Geometry
struct Mesh {
    std::vector<V3> pos;
    std::vector<uint32_t> idx;
};

// Icosphere: icosahedron subdivided `level` times, each triangle replaced by
// its four children in place (keeps triangle order spatially coherent),
// counter-clockwise outward.
Mesh icosphere(uint32_t level)
{
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    Mesh m;
    const float base[12][3] = { { -1, t, 0 }, { 1, t, 0 }, { -1, -t, 0 }, { 1, -t, 0 },
                                { 0, -1, t }, { 0, 1, t }, { 0, -1, -t }, { 0, 1, -t },
                                { t, 0, -1 }, { t, 0, 1 }, { -t, 0, -1 }, { -t, 0, 1 } };
    for (auto& b : base) m.pos.push_back(normalize(V3{ b[0], b[1], b[2] }));
    const uint32_t faces[20][3] = { { 0, 11, 5 }, { 0, 5, 1 }, { 0, 1, 7 }, { 0, 7, 10 }, { 0, 10, 11 },
                                    { 1, 5, 9 }, { 5, 11, 4 }, { 11, 10, 2 }, { 10, 7, 6 }, { 7, 1, 8 },
                                    { 3, 9, 4 }, { 3, 4, 2 }, { 3, 2, 6 }, { 3, 6, 8 }, { 3, 8, 9 },
                                    { 4, 9, 5 }, { 2, 4, 11 }, { 6, 2, 10 }, { 8, 6, 7 }, { 9, 8, 1 } };
    for (auto& f : faces) { m.idx.push_back(f[0]); m.idx.push_back(f[1]); m.idx.push_back(f[2]); }
    for (uint32_t l = 0; l < level; ++l) {
        std::unordered_map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            const uint64_t key = a < b ? (uint64_t(a) << 32 | b) : (uint64_t(b) << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const uint32_t i = static_cast<uint32_t>(m.pos.size());
            m.pos.push_back(normalize((m.pos[a] + m.pos[b]) * 0.5f));
            mid.emplace(key, i);
            return i;
        };
        std::vector<uint32_t> out;
        out.reserve(m.idx.size() * 4);
        for (size_t i = 0; i < m.idx.size(); i += 3) {
            const uint32_t a = m.idx[i], b = m.idx[i + 1], c = m.idx[i + 2];
            const uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const uint32_t tri[12] = { a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca };
            out.insert(out.end(), tri, tri + 12);
        }
        m.idx.swap(out);
    }
    // Make every face counter-clockwise seen from outside.
    for (size_t i = 0; i < m.idx.size(); i += 3) {
        const V3 a = m.pos[m.idx[i]], b = m.pos[m.idx[i + 1]], c = m.pos[m.idx[i + 2]];
        if (dot(cross(b - a, c - a), a + b + c) < 0.0f) std::swap(m.idx[i + 1], m.idx[i + 2]);
    }
    return m;
}

// Unit cube [-0.5, 0.5]^3, 8 vertices, 12 counter-clockwise outward triangles.
Mesh unit_box()
{
    Mesh m;
    for (int i = 0; i < 8; ++i)
        m.pos.push_back(V3{ (i & 1) ? 0.5f : -0.5f, (i & 2) ? 0.5f : -0.5f, (i & 4) ? 0.5f : -0.5f });
    const uint32_t q[6][4] = { { 0, 2, 3, 1 }, { 4, 5, 7, 6 }, { 0, 1, 5, 4 }, { 2, 6, 7, 3 }, { 0, 4, 6, 2 }, { 1, 3, 7, 5 } };
    for (auto& f : q) {
        const uint32_t tri[6] = { f[0], f[1], f[2], f[0], f[2], f[3] };
        m.idx.insert(m.idx.end(), tri, tri + 6);
    }
    for (size_t i = 0; i < m.idx.size(); i += 3) {
        const V3 a = m.pos[m.idx[i]], b = m.pos[m.idx[i + 1]], c = m.pos[m.idx[i + 2]];
        if (dot(cross(b - a, c - a), a + b + c) < 0.0f) std::swap(m.idx[i + 1], m.idx[i + 2]);
    }
    return m;
}

// Lattice value noise: hashed lattice values in [-1,1], smoothstep-trilinear.
struct ValueNoise {
    uint64_t seed;
    float lattice(int x, int y, int z) const
    {
        uint64_t h = seed ^ (uint64_t(uint32_t(x)) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(uint32_t(y)) * 0xC2B2AE3D27D4EB4Full)
                     ^ (uint64_t(uint32_t(z)) * 0x165667B19E3779F9ull);
        uint64_t s = h;
        const uint64_t r = Rng::splitmix(s);
        return static_cast<float>(r >> 40) * (2.0f / 16777216.0f) - 1.0f;
    }
    float operator()(V3 p) const
    {
        const float fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
        const int x = int(fx), y = int(fy), z = int(fz);
        auto sm = [](float t) { return t * t * (3.0f - 2.0f * t); };
        const float u = sm(p.x - fx), v = sm(p.y - fy), w = sm(p.z - fz);
        auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
        const float c00 = lerp(lattice(x, y, z), lattice(x + 1, y, z), u);
        const float c10 = lerp(lattice(x, y + 1, z), lattice(x + 1, y + 1, z), u);
        const float c01 = lerp(lattice(x, y, z + 1), lattice(x + 1, y, z + 1), u);
        const float c11 = lerp(lattice(x, y + 1, z + 1), lattice(x + 1, y + 1, z + 1), u);
        return lerp(lerp(c00, c10, v), lerp(c01, c11, v), w);
    }
    // Three octaves, frequencies 2/4/8 on the unit sphere, amplitudes 1/0.5/0.25, normalized to [-1,1].
    float fbm3(V3 p) const
    {
        return (operator()(p * 2.0f) + 0.5f * operator()(p * 4.0f) + 0.25f * operator()(p * 8.0f)) / 1.75f;
    }
};

// ---------------------------------------------------------------------------
// Meshlet builder.  Greedy growth in the spirit of meshoptimizer's
// buildMeshlets: the next triangle is taken among those adjacent to the last
// added triangle (fewest new vertices, then nearest to the meshlet centroid),
// falling back to the whole meshlet boundary and then to the next unused
// triangle in index order.  Triangle data of each meshlet is padded to a
// multiple of 4 bytes (assets.cpp:339).
struct MeshletData {
    std::vector<uint32_t> vertex_indices;
    std::vector<uint8_t> triangles;
    std::vector<vkgv_meshlet> meshlets;
    std::vector<vkgv_meshlet_cone> cones;
};

void compute_bounds(const V3* pos, MeshletData& md, const vkgv_meshlet& m, vkgv_meshlet& out, vkgv_meshlet_cone& cone)
{
    // AABB exactly as assets.cpp:349-372: min/max over the meshlet's vertices.
    V3 mn = pos[md.vertex_indices[m.vertex_offset]], mx = mn;
    for (uint32_t i = 1; i < m.vertex_count; ++i) {
        const V3 p = pos[md.vertex_indices[m.vertex_offset + i]];
        mn = vmin(mn, p);
        mx = vmax(mx, p);
    }
    const V3 center = (mn + mx) * 0.5f;
    const V3 ext = mx - center;
    out = m;
    out.aabb_center[0] = center.x; out.aabb_center[1] = center.y; out.aabb_center[2] = center.z;
    out.aabb_extents[0] = ext.x; out.aabb_extents[1] = ext.y; out.aabb_extents[2] = ext.z;

    // Normal cone: axis = normalized mean of unit triangle normals; the
    // meshlet is back-facing from every eye inside the cone
    // dot(normalize(apex - eye), axis) >= sin(max angle(normal, axis)),
    // with the apex behind every triangle plane.
    std::vector<V3> nrm, p0;
    V3 sum{ 0, 0, 0 };
    for (uint32_t t = 0; t < m.triangle_count; ++t) {
        const uint8_t* tri = &md.triangles[m.triangle_offset + 3 * t];
        const V3 a = pos[md.vertex_indices[m.vertex_offset + tri[0]]];
        const V3 b = pos[md.vertex_indices[m.vertex_offset + tri[1]]];
        const V3 c = pos[md.vertex_indices[m.vertex_offset + tri[2]]];
        const V3 n = cross(b - a, c - a);
        const float l = length(n);
        if (!(l > 0.0f)) continue;
        const V3 un = n * (1.0f / l);
        nrm.push_back(un);
        p0.push_back(a);
        sum = sum + un;
    }
    cone = vkgv_meshlet_cone{};
    cone.apex[0] = center.x; cone.apex[1] = center.y; cone.apex[2] = center.z;
    cone.axis[0] = 0; cone.axis[1] = 0; cone.axis[2] = 1;
    cone.cutoff = 1.0f;   // never culls
    const float sl = length(sum);
    if (nrm.empty() || !(sl > 0.0f)) return;
    const V3 axis = sum * (1.0f / sl);
    double mindp = 1.0;
    for (const V3& n : nrm) mindp = std::min(mindp, static_cast<double>(dot(n, axis)));
    mindp -= 1e-3;   // angular slack for rounding in the transform and the test
    if (mindp <= 0.1) return;
    double tmax = -1e30;
    for (size_t i = 0; i < nrm.size(); ++i) {
        const double num = dot(nrm[i], center - p0[i]);
        const double den = dot(nrm[i], axis);
        tmax = std::max(tmax, num / den);
    }
    tmax += 1e-4 * (1.0 + length(ext));
    const V3 apex = center - axis * static_cast<float>(tmax);
    cone.apex[0] = apex.x; cone.apex[1] = apex.y; cone.apex[2] = apex.z;
    cone.axis[0] = axis.x; cone.axis[1] = axis.y; cone.axis[2] = axis.z;
    cone.cutoff = static_cast<float>(std::sqrt(1.0 - mindp * mindp) + 1e-4);
}

int build_meshlets(const V3* pos, uint32_t nv, const uint32_t* idx, uint64_t ni, uint32_t maxv, uint32_t maxt,
                   MeshletData& md)
{
    if (ni % 3 || maxv == 0 || maxv > 255 || maxt == 0 || maxt > VKGV_MESHLET_TRIANGLES || (maxt % 4)) {
        g_error = "build_meshlets: invalid limits or index count";
        return -1;
    }
    for (uint64_t i = 0; i < ni; ++i)
        if (idx[i] >= nv) { g_error = "build_meshlets: index out of range"; return -1; }
    const uint64_t nt = ni / 3;
    // vertex -> triangle adjacency (CSR)
    std::vector<uint32_t> off(nv + 1, 0), adj(ni);
    for (uint64_t i = 0; i < ni; ++i) off[idx[i] + 1]++;
    for (uint32_t v = 0; v < nv; ++v) off[v + 1] += off[v];
    {
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (uint64_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) adj[fill[idx[3 * t + k]]++] = static_cast<uint32_t>(t);
    }
    std::vector<uint8_t> used(nt, 0);
    std::vector<uint32_t> live(nv);   // unused triangles incident to each vertex
    for (uint32_t v = 0; v < nv; ++v) live[v] = off[v + 1] - off[v];
    std::vector<int16_t> local(nv, -1);
    std::vector<uint32_t> cur_verts;
    std::vector<uint32_t> cur_tris;
    V3 centroid_sum{ 0, 0, 0 };
    uint64_t seed_cursor = 0;

    auto tri_centroid = [&](uint64_t t) {
        return (pos[idx[3 * t]] + pos[idx[3 * t + 1]] + pos[idx[3 * t + 2]]) * (1.0f / 3.0f);
    };
    auto new_verts = [&](uint64_t t) {
        int n = 0;
        for (int k = 0; k < 3; ++k) n += local[idx[3 * t + k]] < 0;
        return n;
    };
    std::vector<uint32_t> prev_verts;   // vertices of the last finished meshlet (seed search)
    auto flush = [&]() {
        if (cur_tris.empty()) return;
        prev_verts = cur_verts;
        vkgv_meshlet m{};
        m.vertex_offset = static_cast<uint32_t>(md.vertex_indices.size());
        m.triangle_offset = static_cast<uint32_t>(md.triangles.size());
        m.vertex_count = static_cast<uint8_t>(cur_verts.size());
        m.triangle_count = static_cast<uint8_t>(cur_tris.size());
        md.vertex_indices.insert(md.vertex_indices.end(), cur_verts.begin(), cur_verts.end());
        for (uint32_t t : cur_tris)
            for (int k = 0; k < 3; ++k) md.triangles.push_back(static_cast<uint8_t>(local[idx[3 * uint64_t(t) + k]]));
        while (md.triangles.size() % 4) md.triangles.push_back(0);
        md.meshlets.push_back(m);
        for (uint32_t v : cur_verts) local[v] = -1;
        cur_verts.clear();
        cur_tris.clear();
        centroid_sum = V3{ 0, 0, 0 };
    };
    auto add = [&](uint64_t t) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = idx[3 * t + k];
            if (local[v] < 0) {
                local[v] = static_cast<int16_t>(cur_verts.size());
                cur_verts.push_back(v);
            }
        }
        cur_tris.push_back(static_cast<uint32_t>(t));
        used[t] = 1;
        for (int k = 0; k < 3; ++k) --live[idx[3 * t + k]];
        centroid_sum = centroid_sum + tri_centroid(t);
    };

    uint64_t last = UINT64_MAX;
    uint64_t remaining = nt;
    while (remaining) {
        uint64_t best = UINT64_MAX;
        if (last != UINT64_MAX) {
            const V3 cen = centroid_sum * (1.0f / static_cast<float>(cur_tris.size()));
            // Score: fewest new vertices first, then fewest live triangles
            // left on its vertices (finish corners instead of stranding
            // them), then nearest to the meshlet centroid.
            int best_score = 1 << 30;
            float best_d = 0.0f;
            auto consider = [&](uint32_t v) {
                for (uint32_t a = off[v]; a < off[v + 1]; ++a) {
                    const uint32_t t = adj[a];
                    if (used[t]) continue;
                    int lv = 0;
                    for (int k = 0; k < 3; ++k) lv += static_cast<int>(live[idx[3 * uint64_t(t) + k]]);
                    const int score = new_verts(t) * 64 + lv;
                    const V3 d = tri_centroid(t) - cen;
                    const float dd = dot(d, d);
                    if (score < best_score || (score == best_score && dd < best_d)) { best = t; best_score = score; best_d = dd; }
                }
            };
            for (int k = 0; k < 3; ++k) consider(idx[3 * last + k]);
            if (best == UINT64_MAX)
                for (uint32_t v : cur_verts) consider(v);
            if (best != UINT64_MAX
                && (cur_verts.size() + new_verts(best) > maxv || cur_tris.size() + 1 > maxt)) {
                flush();
            }
        }
        if (best == UINT64_MAX || cur_tris.empty()) {
            if (best == UINT64_MAX) {
                if (!cur_tris.empty()) flush();
                // Seed next to the meshlet just finished, at the unused
                // triangle with the fewest unused neighbours, so that holes
                // are filled rather than left as fragments.
                int best_free = 1 << 30;
                for (uint32_t v : prev_verts)
                    for (uint32_t a = off[v]; a < off[v + 1]; ++a) {
                        const uint32_t t = adj[a];
                        if (used[t]) continue;
                        int nfree = 0;
                        for (int k = 0; k < 3; ++k) {
                            const uint32_t w = idx[3 * uint64_t(t) + k];
                            for (uint32_t b = off[w]; b < off[w + 1]; ++b) nfree += !used[adj[b]];
                        }
                        if (nfree < best_free) { best_free = nfree; best = t; }
                    }
                if (best == UINT64_MAX) {
                    while (used[seed_cursor]) ++seed_cursor;
                    best = seed_cursor;
                }
            }
        }
        add(best);
        last = best;
        --remaining;
    }
    flush();

    md.cones.resize(md.meshlets.size());
    for (size_t i = 0; i < md.meshlets.size(); ++i)
        compute_bounds(pos, md, md.meshlets[i], md.meshlets[i], md.cones[i]);
    return 0;
}

// Recomputes AABBs and cones of an existing partition for displaced positions.
void rebound(const V3* pos, MeshletData& md)
{
    for (size_t i = 0; i < md.meshlets.size(); ++i) compute_bounds(pos, md, md.meshlets[i], md.meshlets[i], md.cones[i]);
}

uint8_t to_unorm8(float v) { v = std::min(std::max(v, 0.0f), 1.0f); return static_cast<uint8_t>(v * 255.0f + 0.5f); }

} // namespace

// ---------------------------------------------------------------------------
// Scene
struct vkgv_scene {
    struct MeshStore {
        MeshletData md;
        std::vector<vkgv_vertex> vertices;
    };
    std::vector<std::unique_ptr<MeshStore>> stores;
    std::vector<vkgv_mesh_desc> meshes;
    std::vector<vkgv_primitive_desc> primitives;
    std::vector<vkgv_material> materials;
    std::vector<vkgv_meshlet_draw> draws;
    std::vector<float> transforms;
    std::vector<vkgv_view_desc> views;
    uint64_t triangles = 0;

    uint32_t add_mesh(std::unique_ptr<MeshStore> st)
    {
        stores.push_back(std::move(st));
        return static_cast<uint32_t>(stores.size() - 1);
    }
    void finalize()
    {
        meshes.clear();
        for (auto& st : stores) {
            vkgv_mesh_desc d{};
            d.vertex_indices = st->md.vertex_indices.data();
            d.vertex_index_count = st->md.vertex_indices.size();
            d.triangle_indices = st->md.triangles.data();
            d.triangle_index_bytes = st->md.triangles.size();
            d.vertices = st->vertices.data();
            d.vertex_count = st->vertices.size();
            d.meshlets = st->md.meshlets.data();
            d.meshlet_count = static_cast<uint32_t>(st->md.meshlets.size());
            d.cones = st->md.cones.data();
            meshes.push_back(d);
        }
        triangles = 0;
        for (auto& dr : draws) {
            const auto& p = primitives[dr.primitive_index];
            triangles += stores[p.mesh_index]->md.meshlets[dr.meshlet_index].triangle_count;
        }
    }
    // One draw per meshlet of `prim`, transform index = next transform.
    void add_instance(uint32_t prim, const M4& m)
    {
        const uint32_t ti = static_cast<uint32_t>(transforms.size() / 16);
        transforms.insert(transforms.end(), m.begin(), m.end());
        const auto& st = *stores[primitives[prim].mesh_index];
        for (uint32_t i = 0; i < st.md.meshlets.size(); ++i) draws.push_back({ prim, i, ti });
    }
    uint32_t add_primitive(uint32_t mesh, uint32_t material)
    {
        vkgv_primitive_desc p{};
        p.mesh_index = mesh;
        p.material_index = material;
        // Primitive AABB from the vertex positions (accessor min/max, assets.cpp:302-307).
        const auto& vs = stores[mesh]->vertices;
        V3 mn{ vs[0].position[0], vs[0].position[1], vs[0].position[2] }, mx = mn;
        for (auto& v : vs) {
            const V3 q{ v.position[0], v.position[1], v.position[2] };
            mn = vmin(mn, q);
            mx = vmax(mx, q);
        }
        const V3 c = (mn + mx) * 0.5f, e = mx - c;
        p.aabb_center[0] = c.x; p.aabb_center[1] = c.y; p.aabb_center[2] = c.z;
        p.aabb_extents[0] = e.x; p.aabb_extents[1] = e.y; p.aabb_extents[2] = e.z;
        primitives.push_back(p);
        return static_cast<uint32_t>(primitives.size() - 1);
    }
    uint32_t add_material(float r, float g, float b, float a, uint32_t double_sided = 0)
    {
        vkgv_material m{};
        m.albedo_factor[0] = r; m.albedo_factor[1] = g; m.albedo_factor[2] = b; m.albedo_factor[3] = a;
        m.albedo_index = 0xFFFFFFFFu;
        m.uv_scale[0] = m.uv_scale[1] = 1.0f;
        m.alpha_cutoff = 0.5f;
        m.double_sided = double_sided;
        materials.push_back(m);
        return static_cast<uint32_t>(materials.size() - 1);
    }
};

namespace {

std::unique_ptr<vkgv_scene::MeshStore> make_store(const std::vector<V3>& pos, const std::vector<uint32_t>& idx,
                                                  const MeshletData* partition, const uint8_t (*colors)[4] = nullptr)
{
    auto st = std::make_unique<vkgv_scene::MeshStore>();
    if (partition) {
        st->md = *partition;
        rebound(pos.data(), st->md);
    } else if (build_meshlets(pos.data(), static_cast<uint32_t>(pos.size()), idx.data(), idx.size(),
                              VKGV_MAX_VERTICES, VKGV_MESHLET_TRIANGLES, st->md) != 0) {
        return nullptr;
    }
    st->vertices.resize(pos.size());
    for (size_t i = 0; i < pos.size(); ++i) {
        vkgv_vertex& v = st->vertices[i];
        v.position[0] = pos[i].x; v.position[1] = pos[i].y; v.position[2] = pos[i].z;
        for (int k = 0; k < 4; ++k) v.color[k] = colors ? colors[i][k] : 255;
        const V3 n = normalize(pos[i]);
        v.normal[0] = to_unorm8((n.x + 1.0f) * 0.5f);
        v.normal[1] = to_unorm8((n.y + 1.0f) * 0.5f);
        v.normal[2] = to_unorm8((n.z + 1.0f) * 0.5f);
        v.uv[0] = v.uv[1] = 0;
    }
    return st;
}

M4 translate_scale(V3 t, V3 s, float yaw = 0.0f)
{
    const float q[4] = { 0.0f, std::sin(yaw * 0.5f), 0.0f, std::cos(yaw * 0.5f) };
    const float tt[3] = { t.x, t.y, t.z }, ss[3] = { s.x, s.y, s.z };
    return trs(identity(), tt, q, ss);
}

vkgv_view_desc perspective_view(uint32_t w, uint32_t h, V3 eye, V3 target, float fovy_deg, float n = 0.1f, float f = 1000.0f)
{
    vkgv_view_desc v{};
    v.width = w; v.height = h;
    v.eye[0] = eye.x; v.eye[1] = eye.y; v.eye[2] = eye.z;
    v.target[0] = target.x; v.target[1] = target.y; v.target[2] = target.z;
    v.up[1] = 1.0f;
    v.fovy = fovy_deg * 3.14159265358979f / 180.0f;
    v.znear = n; v.zfar = f;
    return v;
}

// Configuration 1: 256 subdivision-4 icospheres, uniform in a 100 m cube.
void scene_c1(vkgv_scene& s, Rng& rng, uint32_t instances)
{
    const Mesh ico = icosphere(4);
    s.add_mesh(make_store(ico.pos, ico.idx, nullptr));
    s.add_material(1, 1, 1, 1);   // default material (index 0)
    for (uint32_t i = 0; i < instances; ++i) {
        const V3 p{ rng.uniform(-50.0f, 50.0f), rng.uniform(-50.0f, 50.0f), rng.uniform(-50.0f, 50.0f) };
        const float sc = rng.uniform(0.5f, 2.0f);
        const float yaw = rng.uniform(0.0f, 6.2831853f);
        const float r = rng.uniform(), g = rng.uniform(), b = rng.uniform();
        const uint32_t mat = s.add_material(r, g, b, 1.0f);
        const uint32_t prim = s.add_primitive(0, mat);
        s.add_instance(prim, translate_scale(p, V3{ sc, sc, sc }, yaw));
    }
}

// Configuration 2: city of 64x64 boxes (12 m cells, 8 m footprints, heights
// log-uniform 2..80 m) with noise-displaced int16-quantized spheres at the
// street corners of the cells; street-level camera.
int scene_c2(vkgv_scene& s, Rng& rng, uint32_t variant)
{
    const uint32_t level = variant == 1 ? 5 : 4;
    const bool checker = variant != 1;
    const Mesh box = unit_box();
    const uint32_t box_mesh = s.add_mesh(make_store(box.pos, box.idx, nullptr));
    s.add_material(1, 1, 1, 1);
    const Mesh ico = icosphere(level);
    MeshletData partition;
    if (build_meshlets(ico.pos.data(), static_cast<uint32_t>(ico.pos.size()), ico.idx.data(), ico.idx.size(),
                       VKGV_MAX_VERTICES, VKGV_MESHLET_TRIANGLES, partition) != 0)
        return -1;
    const float cell = 12.0f, half_grid = 32.0f * cell;
    struct Sphere { uint32_t i, j; uint64_t seed; };
    std::vector<Sphere> spheres;
    std::vector<float> heights(64 * 64);
    for (uint32_t j = 0; j < 64; ++j)
        for (uint32_t i = 0; i < 64; ++i) {
            heights[j * 64 + i] = 2.0f * std::exp(rng.uniform() * std::log(40.0f));
            if (!checker || ((i + j) & 1) == 0) spheres.push_back({ i, j, (uint64_t(rng.next()) << 32) | rng.next() });
        }
    // Per-sphere meshes: radial 3-octave value noise (amplitude 0.1), then
    // KHR_mesh_quantization int16 normalized positions; the dequantization
    // scale goes into the node transform.
    std::vector<std::unique_ptr<vkgv_scene::MeshStore>> stores(spheres.size());
    std::vector<float> scales(spheres.size());
#pragma omp parallel for schedule(dynamic, 8)
    for (long si = 0; si < (long)spheres.size(); ++si) {
        const ValueNoise noise{ spheres[si].seed };
        std::vector<V3> p(ico.pos.size());
        float m = 0.0f;
        for (size_t v = 0; v < p.size(); ++v) {
            const float r = 1.0f + 0.1f * noise.fbm3(ico.pos[v]);
            p[v] = ico.pos[v] * r;
            m = std::max(m, std::max(std::fabs(p[v].x), std::max(std::fabs(p[v].y), std::fabs(p[v].z))));
        }
        for (auto& q : p) {
            const auto quant = [m](float x) {
                const long c = std::lround(x / m * 32767.0f);
                const float dq = static_cast<float>(std::max(std::min(c, 32767L), -32767L)) / 32767.0f;
                return std::max(dq, -1.0f);
            };
            q = V3{ quant(q.x), quant(q.y), quant(q.z) };
        }
        scales[si] = m;
        stores[si] = make_store(p, ico.idx, &partition);
    }
    Rng box_rng(rng.next());
    for (uint32_t j = 0; j < 64; ++j)
        for (uint32_t i = 0; i < 64; ++i) {
            const float h = heights[j * 64 + i];
            const V3 c{ -half_grid + (float(i) + 0.5f) * cell, h * 0.5f, -half_grid + (float(j) + 0.5f) * cell };
            const float g = 0.35f + 0.4f * box_rng.uniform();
            const uint32_t box_mat = s.add_material(g, g * 0.95f, g * 0.9f, 1.0f);
            const uint32_t box_prim = s.add_primitive(box_mesh, box_mat);
            s.add_instance(box_prim, translate_scale(c, V3{ 8.0f, h, 8.0f }));
        }
    for (size_t si = 0; si < spheres.size(); ++si) {
        const uint32_t mesh = s.add_mesh(std::move(stores[si]));
        const float r = 0.9f;
        const uint32_t mat = s.add_material(rng.uniform(), rng.uniform(), rng.uniform(), 1.0f);
        const uint32_t prim = s.add_primitive(mesh, mat);
        const V3 c{ -half_grid + float(spheres[si].i) * cell + 1.0f, r, -half_grid + float(spheres[si].j) * cell + 1.0f };
        const float sc = r * scales[si];
        s.add_instance(prim, translate_scale(c, V3{ sc, sc, sc }));
    }
    // Street-level camera in the street between rows 31 and 32, looking
    // down it with a slight yaw.
    const V3 eye{ -half_grid + 0.5f * cell, 1.7f, 0.0f };
    const V3 target{ eye.x + 100.0f, 1.7f, 100.0f * std::tan(4.0f * 3.14159265f / 180.0f) };
    s.views.push_back(perspective_view(3840, 2160, eye, target, 60.0f));
    return 0;
}

// Configuration 3: configuration 1's scene seen from 64 views.
void views_c3(vkgv_scene& s, Rng& rng)
{
    const V3 light = normalize(V3{ -0.4f, -1.0f, -0.3f });
    const float cascade[4] = { 15.0f, 30.0f, 60.0f, 120.0f };
    for (int c = 0; c < 4; ++c) {
        vkgv_view_desc v{};
        v.width = v.height = 2048;
        v.orthographic = 1;
        const V3 eye = light * -200.0f;
        v.eye[0] = eye.x; v.eye[1] = eye.y; v.eye[2] = eye.z;
        v.up[0] = 0; v.up[1] = 0; v.up[2] = 1;
        v.znear = 1.0f; v.zfar = 400.0f;
        v.ortho_half_width = v.ortho_half_height = cascade[c];
        s.views.push_back(v);
    }
    const V3 dirs[6] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    const V3 ups[6] = { { 0, 1, 0 }, { 0, 1, 0 }, { 0, 0, 1 }, { 0, 0, 1 }, { 0, 1, 0 }, { 0, 1, 0 } };
    for (int f = 0; f < 6; ++f) {
        vkgv_view_desc v = perspective_view(512, 512, V3{ 0, 0, 0 }, dirs[f], 90.0f);
        v.up[0] = ups[f].x; v.up[1] = ups[f].y; v.up[2] = ups[f].z;
        s.views.push_back(v);
    }
    for (int k = 0; k < 54; ++k) {
        const uint32_t w = 256 + rng.next() % (1920 - 256 + 1);
        const uint32_t h = 256 + rng.next() % (1080 - 256 + 1);
        V3 eye{ rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1) };
        eye = normalize(eye) * rng.uniform(60.0f, 120.0f);
        const V3 tgt{ rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(-20, 20) };
        s.views.push_back(perspective_view(w, h, eye, tgt, rng.uniform(40.0f, 90.0f)));
    }
}

// Configuration 4: one subdivision-6 icosphere (40,962 vertices) with
// per-vertex albedo, eight cameras on a ring at 512x512.
void scene_c4(vkgv_scene& s, Rng& rng)
{
    const Mesh ico = icosphere(6);
    std::vector<std::array<uint8_t, 4>> col(ico.pos.size());
    const ValueNoise noise{ (uint64_t(rng.next()) << 32) | rng.next() };
    for (size_t i = 0; i < col.size(); ++i) {
        const V3 p = ico.pos[i];
        col[i] = { to_unorm8(0.5f + 0.45f * noise(p * 3.0f)), to_unorm8(0.5f + 0.45f * noise(p * 3.0f + V3{ 7, 1, 3 })),
                   to_unorm8(0.5f + 0.45f * noise(p * 3.0f + V3{ 2, 9, 5 })), 255 };
    }
    s.add_mesh(make_store(ico.pos, ico.idx, nullptr, reinterpret_cast<const uint8_t(*)[4]>(col.data())));
    s.add_material(1, 1, 1, 1);
    const uint32_t prim = s.add_primitive(0, 0);
    s.add_instance(prim, identity());
    for (int k = 0; k < 8; ++k) {
        const float a = 6.2831853f * float(k) / 8.0f;
        s.views.push_back(perspective_view(512, 512, V3{ 3.0f * std::cos(a), 0.8f, 3.0f * std::sin(a) }, V3{ 0, 0, 0 }, 45.0f));
    }
}

// Configuration 5: uneven stress.  (200 >> variant) meshes with log-uniform
// meshlet counts in [1, 20k] (terrain-like height-field patches), positions
// fp32 / int16n / uint8n (dequantized as KHR_mesh_quantization), 1-64
// instances each, 16 views of mixed resolution up to 3840x2160.
int scene_c5(vkgv_scene& s, Rng& rng, uint32_t variant)
{
    const uint32_t nmesh = std::max(1u, 200u >> variant);
    s.add_material(1, 1, 1, 1);
    uint64_t draws = 0;
    for (uint32_t mi = 0; mi < nmesh; ++mi) {
        const float lm = rng.uniform() * std::log(20000.0f);
        const uint32_t target_meshlets = std::max(1u, static_cast<uint32_t>(std::exp(lm)));
        // grid of n x n quads, 2 triangles each, about 100 triangles per meshlet
        const uint32_t n = std::max(1u, static_cast<uint32_t>(std::sqrt(double(target_meshlets) * 50.0)));
        std::vector<V3> pos;
        std::vector<uint32_t> idx;
        const ValueNoise noise{ (uint64_t(rng.next()) << 32) | rng.next() };
        for (uint32_t y = 0; y <= n; ++y)
            for (uint32_t x = 0; x <= n; ++x) {
                const float fx = float(x) / float(n) * 2.0f - 1.0f, fz = float(y) / float(n) * 2.0f - 1.0f;
                pos.push_back(V3{ fx, 0.15f * noise(V3{ fx * 4.0f, 0.5f, fz * 4.0f }), fz });
            }
        // Tile-ordered quads (8x4 tiles) so consecutive triangles are spatially local.
        for (uint32_t ty = 0; ty < n; ty += 4)
            for (uint32_t tx = 0; tx < n; tx += 8)
                for (uint32_t y = ty; y < std::min(ty + 4, n); ++y)
                    for (uint32_t x = tx; x < std::min(tx + 8, n); ++x) {
                        const uint32_t a = y * (n + 1) + x, b = a + 1, c = a + n + 1, d = c + 1;
                        const uint32_t t[6] = { a, c, b, b, c, d };
                        idx.insert(idx.end(), t, t + 6);
                    }
        const uint32_t fmt = rng.next() % 3;   // 0 fp32, 1 int16n, 2 uint8n
        if (fmt == 1)
            for (auto& p : pos) {
                auto q = [](float x) { return std::max(static_cast<float>(std::lround(x * 32767.0f)) / 32767.0f, -1.0f); };
                p = V3{ q(p.x), q(p.y), q(p.z) };
            }
        else if (fmt == 2)
            for (auto& p : pos) {
                auto q = [](float x) { return static_cast<float>(std::lround((x * 0.5f + 0.5f) * 255.0f)) / 255.0f * 2.0f - 1.0f; };
                p = V3{ q(p.x), q(p.y), q(p.z) };
            }
        auto st = make_store(pos, idx, nullptr);
        if (!st) return -1;
        const uint32_t mesh = s.add_mesh(std::move(st));
        const uint32_t mat = s.add_material(rng.uniform(), rng.uniform(), rng.uniform(), 1.0f);
        const uint32_t prim = s.add_primitive(mesh, mat);
        const uint32_t inst = 1 + rng.next() % 64;
        const float size = 4.0f + 0.05f * std::sqrt(float(n) * float(n));
        for (uint32_t k = 0; k < inst; ++k) {
            const V3 c{ rng.uniform(-300.0f, 300.0f), rng.uniform(-5.0f, 5.0f), rng.uniform(-300.0f, 300.0f) };
            draws += s.stores[mesh]->md.meshlets.size();
            if (draws >= VKGV_MAX_DRAWS) { g_error = "scene_c5: draw count exceeds 2^25"; return -1; }
            s.add_instance(prim, translate_scale(c, V3{ size, size, size }, rng.uniform(0.0f, 6.2831853f)));
        }
    }
    for (int k = 0; k < 16; ++k) {
        const uint32_t w = k == 0 ? 3840 : 256 + rng.next() % (3840 - 256 + 1);
        const uint32_t h = k == 0 ? 2160 : 256 + rng.next() % (2160 - 256 + 1);
        const float a = rng.uniform(0.0f, 6.2831853f);
        const V3 eye{ 350.0f * std::cos(a), rng.uniform(20.0f, 120.0f), 350.0f * std::sin(a) };
        s.views.push_back(perspective_view(w, h, eye, V3{ 0, 0, 0 }, rng.uniform(40.0f, 75.0f)));
    }
    return 0;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

void vkgv_camera_from_view(const vkgv_view_desc* view, vkgv_camera* out)
{
    const M4 vp = view_projection(*view);
    std::memcpy(out->view_projection, vp.data(), 64);
    std::memcpy(out->occlusion_view_projection, vp.data(), 64);
    std::memcpy(out->prev_view_projection, vp.data(), 64);
    std::memcpy(out->prev_occlusion_view_projection, vp.data(), 64);
    frustum_planes(vp, out->frustum);
}

void vkgv_camera_advance(const vkgv_camera* prev, const vkgv_view_desc* view, vkgv_camera* out)
{
    vkgv_camera c{};
    std::memcpy(c.prev_view_projection, prev->view_projection, 64);
    std::memcpy(c.prev_occlusion_view_projection, prev->occlusion_view_projection, 64);
    const M4 vp = view_projection(*view);
    std::memcpy(c.view_projection, vp.data(), 64);
    std::memcpy(c.occlusion_view_projection, vp.data(), 64);
    frustum_planes(vp, c.frustum);
    *out = c;
}

int vkgv_build_meshlets(const float* positions, uint32_t vertex_count, const uint32_t* indices, uint64_t index_count,
                        uint32_t max_vertices, uint32_t max_triangles, vkgv_meshlet_build* out)
{
    if (!positions || !indices || !out) { g_error = "vkgv_build_meshlets: null argument"; return -1; }
    std::vector<V3> pos(vertex_count);
    for (uint32_t i = 0; i < vertex_count; ++i) pos[i] = V3{ positions[3 * i], positions[3 * i + 1], positions[3 * i + 2] };
    MeshletData md;
    if (build_meshlets(pos.data(), vertex_count, indices, index_count, max_vertices, max_triangles, md) != 0) return -1;
    auto dup = [](const void* src, size_t bytes) {
        void* p = std::malloc(bytes ? bytes : 1);
        if (bytes) std::memcpy(p, src, bytes);
        return p;
    };
    out->vertex_indices = static_cast<uint32_t*>(dup(md.vertex_indices.data(), md.vertex_indices.size() * 4));
    out->vertex_index_count = md.vertex_indices.size();
    out->triangle_indices = static_cast<uint8_t*>(dup(md.triangles.data(), md.triangles.size()));
    out->triangle_index_bytes = md.triangles.size();
    out->meshlets = static_cast<vkgv_meshlet*>(dup(md.meshlets.data(), md.meshlets.size() * sizeof(vkgv_meshlet)));
    out->meshlet_count = static_cast<uint32_t>(md.meshlets.size());
    out->cones = static_cast<vkgv_meshlet_cone*>(dup(md.cones.data(), md.cones.size() * sizeof(vkgv_meshlet_cone)));
    return 0;
}

void vkgv_meshlet_build_free(vkgv_meshlet_build* b)
{
    if (!b) return;
    std::free(b->vertex_indices);
    std::free(b->triangle_indices);
    std::free(b->meshlets);
    std::free(b->cones);
    std::memset(b, 0, sizeof *b);
}

vkgv_scene* vkgv_scene_generate(uint32_t config, uint64_t seed, uint32_t variant)
{
    auto s = std::make_unique<vkgv_scene>();
    Rng rng(seed);
    int rc = 0;
    switch (config) {
    case 1:
        scene_c1(*s, rng, variant ? variant : 256);
        s->views.push_back(perspective_view(1280, 720, V3{ 0, 20, 110 }, V3{ 0, 0, 0 }, 60.0f));
        break;
    case 2: rc = scene_c2(*s, rng, variant); break;
    case 3:
        scene_c1(*s, rng, 256);
        views_c3(*s, rng);
        break;
    case 4: scene_c4(*s, rng); break;
    case 5: rc = scene_c5(*s, rng, variant); break;
    default: g_error = "vkgv_scene_generate: unknown configuration"; return nullptr;
    }
    if (rc != 0) return nullptr;
    for (auto& st : s->stores)
        if (!st) { g_error = "vkgv_scene_generate: meshlet build failed: " + g_error; return nullptr; }
    s->finalize();
    return s.release();
}

void vkgv_scene_free(vkgv_scene* s) { delete s; }
const char* vkgv_scene_error(void) { return g_error.c_str(); }

uint32_t vkgv_scene_mesh_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->meshes.size()); }
const vkgv_mesh_desc* vkgv_scene_meshes(const vkgv_scene* s) { return s->meshes.data(); }
uint32_t vkgv_scene_primitive_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->primitives.size()); }
const vkgv_primitive_desc* vkgv_scene_primitives(const vkgv_scene* s) { return s->primitives.data(); }
uint32_t vkgv_scene_material_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->materials.size()); }
const vkgv_material* vkgv_scene_materials(const vkgv_scene* s) { return s->materials.data(); }
uint32_t vkgv_scene_draw_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->draws.size()); }
const vkgv_meshlet_draw* vkgv_scene_draws(const vkgv_scene* s) { return s->draws.data(); }
uint32_t vkgv_scene_transform_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->transforms.size() / 16); }
const float* vkgv_scene_transforms(const vkgv_scene* s) { return s->transforms.data(); }
uint32_t vkgv_scene_view_count(const vkgv_scene* s) { return static_cast<uint32_t>(s->views.size()); }
const vkgv_view_desc* vkgv_scene_views(const vkgv_scene* s) { return s->views.data(); }
uint64_t vkgv_scene_triangle_count(const vkgv_scene* s) { return s->triangles; }

int vkgv_world_build(const vkgv_node* nodes, uint32_t node_count, const uint32_t* children, const uint32_t* roots,
                     uint32_t root_count, const uint32_t* mesh_primitives, const uint32_t* prim_list,
                     const uint32_t* meshlet_counts, vkgv_meshlet_draw* draws, uint32_t* draw_count,
                     float* transforms, uint32_t* transform_count)
{
    if (!nodes || !roots || !draw_count || !transform_count) { g_error = "vkgv_world_build: null argument"; return -1; }
    uint64_t nd = 0;
    uint32_t nt = 0;
    struct Frame { uint32_t node; M4 parent; };
    std::vector<Frame> stack;
    // Depth-first, children in order (World::iterateNode recursion order).
    for (uint32_t r = 0; r < root_count; ++r) {
        if (roots[r] >= node_count) { g_error = "vkgv_world_build: root out of range"; return -1; }
        stack.push_back({ roots[r], identity() });
        while (!stack.empty()) {
            Frame f = stack.back();
            stack.pop_back();
            const vkgv_node& n = nodes[f.node];
            const M4 m = trs(f.parent, n.translation, n.rotation, n.scale);
            if (n.mesh >= 0) {
                const uint32_t ti = nt++;
                if (transforms) std::memcpy(transforms + 16ull * ti, m.data(), 64);
                const uint32_t first = mesh_primitives[2 * n.mesh], count = mesh_primitives[2 * n.mesh + 1];
                for (uint32_t k = 0; k < count; ++k) {
                    const uint32_t prim = prim_list[first + k];
                    for (uint32_t i = 0; i < meshlet_counts[prim]; ++i) {
                        if (draws) draws[nd] = vkgv_meshlet_draw{ prim, i, ti };
                        if (++nd > VKGV_MAX_DRAWS) { g_error = "vkgv_world_build: more than 2^25 draws"; return -4; }
                    }
                }
            }
            for (uint32_t c = n.child_count; c-- > 0;) {
                const uint32_t child = children[n.first_child + c];
                if (child >= node_count) { g_error = "vkgv_world_build: child out of range"; return -1; }
                stack.push_back({ child, m });
            }
        }
    }
    *draw_count = static_cast<uint32_t>(nd);
    *transform_count = nt;
    return 0;
}

} // extern "C"
