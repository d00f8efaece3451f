// LibTorch bindings of the differentiable form: rasterize (no gradient),
// interpolate and antialias as torch::autograd::Function's whose backward
// runs the hand-written kernels of diff.hip.
#include <torch/extension.h>
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

struct InterpolateFn : public torch::autograd::Function<InterpolateFn> {
    static Tensor forward(AutogradContext* ctx, Tensor clip, Tensor attr, Tensor tris, Tensor ids)
    {
        const Shape s = shape_of(clip, tris, ids, attr.dim() == 2 ? (int)attr.size(1) : 0);
        check_real(attr, "attr", clip);
        TORCH_CHECK(attr.dim() == 2 && attr.size(0) == clip.size(2), "attr must be [V, A]");
        Tensor out = torch::empty({ s.C, s.H, s.W, s.A }, attr.options());
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::interp_fwd(clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(), ids.data_ptr<int>(), s.C,
                                  s.I, s.V, s.T, s.H, s.W, s.A, out.data_ptr<R>(), stream());
        });
        ctx->save_for_backward({ clip, attr, tris, ids });
        return out;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &clip = saved[0], &attr = saved[1], &tris = saved[2], &ids = saved[3];
        const Shape s = shape_of(clip, tris, ids, (int)attr.size(1));
        Tensor g_out = grads[0].to(attr.scalar_type()).contiguous();
        Tensor g_clip = torch::zeros_like(clip), g_attr = torch::zeros_like(attr);
        dispatch_real(clip, attr, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::interp_bwd(clip.data_ptr<C>(), attr.data_ptr<R>(), tris.data_ptr<int>(), ids.data_ptr<int>(), s.C,
                                  s.I, s.V, s.T, s.H, s.W, s.A, g_out.data_ptr<R>(), g_clip.data_ptr<C>(),
                                  g_attr.data_ptr<R>(), stream());
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
        dispatch_real(clip, color, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::aa_fwd(color.data_ptr<R>(), clip.data_ptr<C>(), tris.data_ptr<int>(), ids.data_ptr<int>(),
                              depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A, out.data_ptr<R>(), stream());
        });
        ctx->save_for_backward({ color, clip, tris, ids, depth });
        return out;
    }
    static variable_list backward(AutogradContext* ctx, variable_list grads)
    {
        const auto saved = ctx->get_saved_variables();
        const Tensor &color = saved[0], &clip = saved[1], &tris = saved[2], &ids = saved[3], &depth = saved[4];
        const Shape s = shape_of(clip, tris, ids, (int)color.size(3));
        Tensor g_out = grads[0].to(color.scalar_type()).contiguous();
        Tensor g_color = torch::empty_like(color), g_clip = torch::zeros_like(clip);
        dispatch_real(clip, color, [&](auto cc, auto r) {
            using C = decltype(cc);
            using R = decltype(r);
            vkgv_diff::aa_bwd(color.data_ptr<R>(), clip.data_ptr<C>(), tris.data_ptr<int>(), ids.data_ptr<int>(),
                              depth.data_ptr<float>(), s.C, s.I, s.V, s.T, s.H, s.W, s.A, g_out.data_ptr<R>(),
                              g_color.data_ptr<R>(), g_clip.data_ptr<C>(), stream());
        });
        return { g_color, g_clip, Tensor(), Tensor(), Tensor() };
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

std::vector<Tensor> rasterize(Tensor clip, Tensor tris, int64_t H, int64_t W, bool cull_back)
{
    check_f32(clip, "clip");
    check_i32(tris, "tris");
    TORCH_CHECK(clip.dim() == 4 && clip.size(3) == 4, "clip must be [C, I, V, 4]");
    TORCH_CHECK(H > 0 && W > 0 && H <= 16384 && W <= 16384, "resolution 1..16384");
    TORCH_CHECK(clip.size(1) * tris.size(0) < (int64_t)0x7FFFFFFF, "instances x triangles must fit the 31-bit id");
    if (tris.numel()) {
        const auto mx = tris.max().item<int>(), mn = tris.min().item<int>();
        TORCH_CHECK(mn >= 0 && mx < clip.size(2), "triangle vertex index out of range");
    }
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

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "Differentiable form of the vk_gltf_viewer frame on MI355X (HIP kernels)";
    m.def("transform", &transform, "clip positions [C, I, V, 4] from positions [V, 3] and VP M matrices [C, I, 4, 4]");
    m.def("rasterize", &rasterize, "triangle raster -> (ids, depth); no gradient");
    m.def("interpolate", &interpolate, "perspective-correct interpolation of per-vertex attributes");
    m.def("antialias", &antialias, "analytic edge antialiasing");
}
