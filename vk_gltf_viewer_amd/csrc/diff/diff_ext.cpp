// LibTorch bindings of the differentiable form: rasterize (no gradient),
// interpolate and antialias as torch::autograd::Function's whose backward
// runs the hand-written kernels of diff.hip.
#include <torch/extension.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include <c10/hip/HIPStream.h>

#include "diff.hpp"

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

namespace {

hipStream_t stream() { return c10::hip::getCurrentHIPStream().stream(); }

void check_f32(const Tensor& t, const char* name)
{
    TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat32, name, " must be float32");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

// Clip positions: float32 or float64.  Attributes, images and their
// gradients: the clip positions' type, or float32 with float64 clip
// positions (the offsets the results are ill-conditioned in are then formed
// in float64, the rest computed in float32).
void check_real(const Tensor& t, const char* name, const Tensor& clip)
{
    TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kFloat32 || t.scalar_type() == torch::kFloat64, name,
                " must be float32 or float64");
    TORCH_CHECK(t.scalar_type() == clip.scalar_type() || t.scalar_type() == torch::kFloat32, name,
                " must have the clip positions' type or be float32");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

void check_i32(const Tensor& t, const char* name)
{
    TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
    TORCH_CHECK(t.scalar_type() == torch::kInt32, name, " must be int32");
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

struct Shape {
    int C, I, V, T, H, W, A;
};

// clip [C, I, V, 4], tris [T, 3], ids [C, H, W], attr [V, A] (A <= 8)
Shape shape_of(const Tensor& clip, const Tensor& tris, const Tensor& ids, int A)
{
    check_real(clip, "clip", clip);
    check_i32(tris, "tris");
    check_i32(ids, "ids");
    TORCH_CHECK(clip.dim() == 4 && clip.size(3) == 4, "clip must be [C, I, V, 4]");
    TORCH_CHECK(tris.dim() == 2 && tris.size(1) == 3, "tris must be [T, 3]");
    TORCH_CHECK(ids.dim() == 3 && ids.size(0) == clip.size(0), "ids must be [C, H, W]");
    TORCH_CHECK(A >= 1 && A <= 8, "1..8 attribute channels");
    const int64_t ntri = clip.size(1) * tris.size(0);
    TORCH_CHECK(ntri < (int64_t)0x7FFFFFFF, "instances x triangles must fit the 31-bit id");
    return Shape{ (int)clip.size(0), (int)clip.size(1), (int)clip.size(2), (int)tris.size(0), (int)ids.size(1),
                  (int)ids.size(2), A };
}

// f(C(), R()) with C the clip positions' type and R the attributes' type
template <typename F> void dispatch_real(const Tensor& clip, const Tensor& attr, F&& f)
{
    if (clip.scalar_type() == torch::kFloat64) {
        if (attr.scalar_type() == torch::kFloat64) f(double(), double());
        else f(double(), float());
    } else {
        f(float(), float());
    }
}

// The vertex adjacency of a mesh for the backward gathers, built on the
// device and kept for the last few meshes: an entry holds a weak reference
// to the triangle tensor's storage and its version counter, so a freed or
// modified mesh is rebuilt.
struct AdjEntry {
    c10::weak_intrusive_ptr<c10::StorageImpl> storage;
    int64_t offset, version, T, V;
    c10::Device device;
    Tensor off, adj;
};
// Never destroyed: its device tensors must not be freed after the runtime at exit.
std::vector<AdjEntry>& adj_cache()
{
    static auto* c = new std::vector<AdjEntry>();
    return *c;
}

std::pair<Tensor, Tensor> adjacency(const Tensor& tris, int V)
{
    static std::mutex mu;   // backwards of several devices run on their own threads
    std::lock_guard<std::mutex> lock(mu);
    auto& cache = adj_cache();
    const int64_t T = tris.size(0);
    for (size_t k = 0; k < cache.size(); ++k) {
        AdjEntry& e = cache[k];
        auto st = e.storage.lock();
        if (st.get() == tris.storage().unsafeGetStorageImpl() && e.offset == tris.storage_offset() &&
            e.version == tris._version() && e.T == T && e.V == V && e.device == tris.device()) {
            std::pair<Tensor, Tensor> r{ e.off, e.adj };
            if (k) std::rotate(cache.begin(), cache.begin() + k, cache.begin() + k + 1);
            return r;
        }
    }
    auto opt = tris.options().dtype(torch::kInt32);
    Tensor off = torch::empty({ 2 * (int64_t)V + 1 }, opt), adj = torch::empty({ std::max<int64_t>(3 * T, 1) }, opt);
    vkgv_diff::mesh_adjacency(tris.data_ptr<int>(), (int)T, V, off.data_ptr<int>(), adj.data_ptr<int>(), stream());
    cache.insert(cache.begin(), AdjEntry{ tris.storage().getWeakStorageImpl(), tris.storage_offset(), tris._version(), T,
                                          V, tris.device(), off, adj });
    if (cache.size() > 4) cache.pop_back();
    return { off, adj };
}

Tensor stray_flags(const Shape& s, const Tensor& like)
{
    return torch::empty({ (int64_t)vkgv_diff::stray_flag_count(s.C, s.H, s.W) }, like.options().dtype(torch::kInt32));
}

Tensor tri_boxes(const Tensor& clip, const Tensor& tris, const Shape& s)
{
    Tensor boxes = torch::empty({ std::max<int64_t>((int64_t)vkgv_diff::tri_boxes_words(s.C, s.I, s.T), 1) },
                                clip.options().dtype(torch::kInt64));
    auto* b = reinterpret_cast<unsigned long long*>(boxes.data_ptr<int64_t>());
    if (clip.scalar_type() == torch::kFloat64)
        vkgv_diff::tri_boxes(clip.data_ptr<double>(), tris.data_ptr<int>(), s.C, s.I, s.V, s.T, s.H, s.W, b, stream());
    else
        vkgv_diff::tri_boxes(clip.data_ptr<float>(), tris.data_ptr<int>(), s.C, s.I, s.V, s.T, s.H, s.W, b, stream());
    return boxes;
}

unsigned long long* box_ptr(const Tensor& boxes)
{
    return reinterpret_cast<unsigned long long*>(boxes.data_ptr<int64_t>());
}

const unsigned* flag_ptr(const Tensor& flags) { return reinterpret_cast<const unsigned*>(flags.data_ptr<int>()); }

void check_gather(const Shape& s)
{
    TORCH_CHECK(((int64_t)s.V + 255) / 256 * s.C * s.I < ((int64_t)1 << 31), "too many camera-instance-vertex triples");
}

struct InterpolateFn : public torch::autograd::Function<InterpolateFn> {
    static Tensor forward(AutogradContext* ctx, Tensor clip, Tensor attr, Tensor tris, Tensor ids)
    {
        const Shape s = shape_of(clip, tris, ids, attr.dim() == 2 ? (int)attr.size(1) : 0);
        check_real(attr, "attr", clip);
        TORCH_CHECK(attr.dim() == 2 && attr.size(0) == clip.size(2), "attr must be [V, A]");
        Tensor out = torch::empty({ s.C, s.H, s.W, s.A }, attr.options());
        Tensor flags = stray_flags(s, ids), boxes = tri_boxes(clip, tris, s);
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::interp_fwd(clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(), ids.data_ptr<int>(), s.C,
                                  s.I, s.V, s.T, s.H, s.W, s.A, out.data_ptr<R>(), box_ptr(boxes),
                                  reinterpret_cast<unsigned*>(flags.data_ptr<int>()), stream());
        });
        ctx->save_for_backward({ clip, attr, tris, ids, flags, boxes });
        return out;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &clip = saved[0], &attr = saved[1], &tris = saved[2], &ids = saved[3], &flags = saved[4];
        const Tensor& boxes = saved[5];
        const Shape s = shape_of(clip, tris, ids, (int)attr.size(1));
        check_gather(s);
        Tensor g_out = grads[0].to(attr.scalar_type()).contiguous();
        Tensor g_clip = torch::empty_like(clip), g_attr = torch::empty_like(attr);
        Tensor part = torch::empty({ std::max<int64_t>((int64_t)s.C * s.I * s.V * s.A, 1) }, attr.options());
        Tensor contrib = torch::empty({ std::max<int64_t>((int64_t)s.C * s.H * s.W * 3 * (3 + s.A), 1) }, attr.options());
        const auto adj = adjacency(tris, s.V);
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::interp_bwd(clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(), ids.data_ptr<int>(), s.C,
                                  s.I, s.V, s.T, s.H, s.W, s.A, box_ptr(boxes), flag_ptr(flags),
                                  adj.first.data_ptr<int>(), adj.second.data_ptr<int>(), g_out.data_ptr<R>(),
                                  contrib.data_ptr<R>(), part.data_ptr<R>(), g_clip.data_ptr<C>(), g_attr.data_ptr<R>(),
                                  stream());
        });
        return { g_clip, g_attr, Tensor(), Tensor() };
    }
};

struct AntialiasFn : public torch::autograd::Function<AntialiasFn> {
    static Tensor forward(AutogradContext* ctx, Tensor color, Tensor clip, Tensor tris, Tensor ids, Tensor depth)
    {
        check_f32(depth, "depth");
        const Shape s = shape_of(clip, tris, ids, color.dim() == 4 ? (int)color.size(3) : 0);
        check_real(color, "color", clip);
        TORCH_CHECK(color.dim() == 4 && color.size(0) == s.C && color.size(1) == s.H && color.size(2) == s.W,
                    "color must be [C, H, W, A]");
        TORCH_CHECK(depth.sizes() == ids.sizes(), "depth must match ids");
        Tensor out = torch::empty_like(color);
        Tensor flags = stray_flags(s, ids), boxes = tri_boxes(clip, tris, s);
        dispatch_real(clip, color, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::aa_fwd(color.data_ptr<R>(), clip.data_ptr<C>(), tris.data_ptr<int>(), ids.data_ptr<int>(),
                              depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A, out.data_ptr<R>(),
                              box_ptr(boxes), reinterpret_cast<unsigned*>(flags.data_ptr<int>()), stream());
        });
        ctx->save_for_backward({ color, clip, tris, ids, depth, flags, boxes });
        return out;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &color = saved[0], &clip = saved[1], &tris = saved[2], &ids = saved[3], &depth = saved[4];
        const Tensor &flags = saved[5], &boxes = saved[6];
        const Shape s = shape_of(clip, tris, ids, (int)color.size(3));
        check_gather(s);
        Tensor g_out = grads[0].to(color.scalar_type()).contiguous();
        Tensor g_color = torch::empty_like(color), g_clip = torch::empty_like(clip);
        Tensor contrib = torch::empty({ std::max<int64_t>((int64_t)s.C * s.H * s.W * 9, 1) }, color.options());
        const auto adj = adjacency(tris, s.V);
        dispatch_real(clip, color, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::aa_bwd(color.data_ptr<R>(), clip.data_ptr<C>(), tris.data_ptr<int>(), ids.data_ptr<int>(),
                              depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A, box_ptr(boxes),
                              flag_ptr(flags), adj.first.data_ptr<int>(), adj.second.data_ptr<int>(),
                              g_out.data_ptr<R>(), contrib.data_ptr<R>(), g_color.data_ptr<R>(), g_clip.data_ptr<C>(),
                              stream());
        });
        return { g_color, g_clip, Tensor(), Tensor(), Tensor() };
    }
};

// antialias(interpolate(clip, attr, tris, ids), clip, tris, ids, depth) as one
// function: the same forward, and one backward whose vertex pass sums both
// stages' terms into a single clip gradient.
struct ShadeFn : public torch::autograd::Function<ShadeFn> {
    static Tensor forward(AutogradContext* ctx, Tensor clip, Tensor attr, Tensor tris, Tensor ids, Tensor depth)
    {
        check_f32(depth, "depth");
        const Shape s = shape_of(clip, tris, ids, attr.dim() == 2 ? (int)attr.size(1) : 0);
        check_real(attr, "attr", clip);
        TORCH_CHECK(attr.dim() == 2 && attr.size(0) == clip.size(2), "attr must be [V, A]");
        TORCH_CHECK(depth.sizes() == ids.sizes(), "depth must match ids");
        Tensor color = torch::empty({ s.C, s.H, s.W, s.A }, attr.options());
        Tensor out = torch::empty_like(color);
        Tensor flags = stray_flags(s, ids), boxes = tri_boxes(clip, tris, s);
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::interp_fwd(clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(), ids.data_ptr<int>(), s.C,
                                  s.I, s.V, s.T, s.H, s.W, s.A, color.data_ptr<R>(), box_ptr(boxes),
                                  reinterpret_cast<unsigned*>(flags.data_ptr<int>()), stream());
            vkgv_diff::aa_fwd(color.data_ptr<R>(), clip.data_ptr<C>(), tris.data_ptr<int>(), ids.data_ptr<int>(),
                              depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A, out.data_ptr<R>(),
                              box_ptr(boxes), nullptr, stream());
        });
        ctx->save_for_backward({ clip, attr, tris, ids, depth, color, flags, boxes });
        return out;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &clip = saved[0], &attr = saved[1], &tris = saved[2], &ids = saved[3], &depth = saved[4];
        const Tensor &color = saved[5], &flags = saved[6], &boxes = saved[7];
        const Shape s = shape_of(clip, tris, ids, (int)attr.size(1));
        check_gather(s);
        Tensor g_out = grads[0].to(attr.scalar_type()).contiguous();
        Tensor g_clip = torch::empty_like(clip), g_attr = torch::empty_like(attr), g_color = torch::empty_like(color);
        const int64_t npx = (int64_t)s.C * s.H * s.W;
        Tensor part = torch::empty({ std::max<int64_t>((int64_t)s.C * s.I * s.V * s.A, 1) }, attr.options());
        Tensor contrib_i = torch::empty({ std::max<int64_t>(npx * 3 * (3 + s.A), 1) }, attr.options());
        Tensor contrib_a = torch::empty({ std::max<int64_t>(npx * 9, 1) }, attr.options());
        const auto adj = adjacency(tris, s.V);
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::shade_bwd(color.data_ptr<R>(), clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(),
                                 ids.data_ptr<int>(), depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A,
                                 box_ptr(boxes), flag_ptr(flags), adj.first.data_ptr<int>(), adj.second.data_ptr<int>(),
                                 g_out.data_ptr<R>(), g_color.data_ptr<R>(), contrib_a.data_ptr<R>(),
                                 contrib_i.data_ptr<R>(), part.data_ptr<R>(), g_clip.data_ptr<C>(),
                                 g_attr.data_ptr<R>(), stream());
        });
        return { g_clip, g_attr, Tensor(), Tensor(), Tensor() };
    }
};

// clip [C, I, V, 4] = A [C, I, 4, 4] (pos [V, 3], 1)
struct TransformFn : public torch::autograd::Function<TransformFn> {
    static Tensor forward(AutogradContext* ctx, Tensor pos, Tensor A)
    {
        check_real(pos, "pos", pos);
        check_real(A, "A", pos);
        TORCH_CHECK(A.scalar_type() == pos.scalar_type(), "A must have the positions' type");
        TORCH_CHECK(pos.dim() == 2 && pos.size(1) == 3, "pos must be [V, 3]");
        TORCH_CHECK(A.dim() == 4 && A.size(2) == 4 && A.size(3) == 4, "A must be [C, I, 4, 4]");
        TORCH_CHECK(A.size(0) * A.size(1) < 65536, "at most 65535 camera-instance pairs");
        TORCH_CHECK(pos.size(0) < (int64_t)0x7FFFFFFF, "too many vertices");
        const int V = (int)pos.size(0), CI = (int)(A.size(0) * A.size(1));
        Tensor clip = torch::empty({ A.size(0), A.size(1), V, 4 }, pos.options());
        if (pos.scalar_type() == torch::kFloat64)
            vkgv_diff::transform_fwd(pos.data_ptr<double>(), A.data_ptr<double>(), CI, V, clip.data_ptr<double>(), stream());
        else
            vkgv_diff::transform_fwd(pos.data_ptr<float>(), A.data_ptr<float>(), CI, V, clip.data_ptr<float>(), stream());
        ctx->save_for_backward({ pos, A });
        return clip;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &pos = saved[0], &A = saved[1];
        Tensor g = grads[0].to(pos.scalar_type()).contiguous();
        const int V = (int)pos.size(0), CI = (int)(A.size(0) * A.size(1));
        Tensor g_pos = torch::empty_like(pos), g_A = torch::zeros_like(A);
        if (pos.scalar_type() == torch::kFloat64)
            vkgv_diff::transform_bwd(g.data_ptr<double>(), pos.data_ptr<double>(), A.data_ptr<double>(), CI, V,
                                     g_pos.data_ptr<double>(), g_A.data_ptr<double>(), stream());
        else
            vkgv_diff::transform_bwd(g.data_ptr<float>(), pos.data_ptr<float>(), A.data_ptr<float>(), CI, V,
                                     g_pos.data_ptr<float>(), g_A.data_ptr<float>(), stream());
        return { g_pos, g_A };
    }
};

Tensor transform(Tensor pos, Tensor A) { return TransformFn::apply(pos, A); }

// Vertex indices of tris within 0..V-1 (else an error).  The check reads a
// minimum and maximum back to the host, a synchronization, so a mesh that
// passed is remembered (storage, offset, version and V, like the adjacency)
// and not checked again.
void check_vertex_range(const Tensor& tris, int64_t V)
{
    struct Checked {
        c10::weak_intrusive_ptr<c10::StorageImpl> storage;
        int64_t offset, version, numel, V;
    };
    static auto* seen = new std::vector<Checked>();   // never destroyed, like the adjacency cache
    static std::mutex mu;
    if (!tris.numel()) return;
    std::lock_guard<std::mutex> lock(mu);
    for (const Checked& e : *seen) {
        auto st = e.storage.lock();
        if (st.get() == tris.storage().unsafeGetStorageImpl() && e.offset == tris.storage_offset() &&
            e.version == tris._version() && e.numel == tris.numel() && e.V == V)
            return;
    }
    const auto mm = torch::aminmax(tris);
    const Tensor both = torch::stack({ std::get<0>(mm), std::get<1>(mm) }).cpu();
    const int mn = both[0].item<int>(), mx = both[1].item<int>();
    TORCH_CHECK(mn >= 0 && mx < V, "triangle vertex index out of range");
    seen->insert(seen->begin(), Checked{ tris.storage().getWeakStorageImpl(), tris.storage_offset(), tris._version(),
                                         tris.numel(), V });
    if (seen->size() > 8) seen->pop_back();
}

std::vector<Tensor> rasterize(Tensor clip, Tensor tris, int64_t H, int64_t W, bool cull_back)
{
    check_f32(clip, "clip");
    check_i32(tris, "tris");
    TORCH_CHECK(clip.dim() == 4 && clip.size(3) == 4, "clip must be [C, I, V, 4]");
    TORCH_CHECK(H > 0 && W > 0 && H <= 16384 && W <= 16384, "resolution 1..16384");
    TORCH_CHECK(clip.size(1) * tris.size(0) < (int64_t)0x7FFFFFFF, "instances x triangles must fit the 31-bit id");
    check_vertex_range(tris, clip.size(2));
    const int C = (int)clip.size(0);
    auto opt = clip.options();
    Tensor keys = torch::full({ C, H, W }, -1, opt.dtype(torch::kInt64));
    Tensor ids = torch::empty({ C, H, W }, opt.dtype(torch::kInt32));
    Tensor depth = torch::empty({ C, H, W }, opt);
    vkgv_diff::raster_tris(clip.data_ptr<float>(), tris.data_ptr<int>(), C, (int)clip.size(1), (int)clip.size(2),
                           (int)tris.size(0), (int)H, (int)W, cull_back ? 1 : 0,
                           reinterpret_cast<unsigned long long*>(keys.data_ptr<int64_t>()), ids.data_ptr<int>(),
                           depth.data_ptr<float>(), stream());
    return { ids, depth };
}

Tensor interpolate(Tensor clip, Tensor attr, Tensor tris, Tensor ids)
{
    return InterpolateFn::apply(clip, attr, tris, ids);
}

Tensor antialias(Tensor color, Tensor clip, Tensor tris, Tensor ids, Tensor depth)
{
    return AntialiasFn::apply(color, clip, tris, ids, depth);
}

Tensor shade(Tensor clip, Tensor attr, Tensor tris, Tensor ids, Tensor depth)
{
    return ShadeFn::apply(clip, attr, tris, ids, depth);
}

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "Differentiable form of the vk_gltf_viewer frame on MI355X (HIP kernels)";
    m.def("transform", &transform, "clip positions [C, I, V, 4] from positions [V, 3] and VP M matrices [C, I, 4, 4]");
    m.def("rasterize", &rasterize, "triangle raster -> (ids, depth); no gradient");
    m.def("interpolate", &interpolate, "perspective-correct interpolation of per-vertex attributes");
    m.def("antialias", &antialias, "analytic edge antialiasing");
    m.def("shade", &shade, "antialias(interpolate(clip, attr, tris, ids), clip, tris, ids, depth) with one backward");
}
