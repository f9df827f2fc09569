// libtorch binding of the differentiable resolve: a torch::autograd::Function
// whose forward is lh_resolve_ex and whose backward is the hand-written adjoint
// lh_resolve_backward (include/lh/lh.h).  No kernels live here; every call goes
// through the C ABI of liblh.so.
//
// Inputs of the Function (float32 CUDA tensors on the view's device):
//   geometry       the scene's geometry arena viewed as float words (it IS the
//                  scene memory, so updating it moves the vertices the next
//                  resolve reads); its gradient mirrors the arena;
//   base_color     [num_mesh_infos, 3];  roughness, metallic  [num_mesh_infos];
//   sun_dir, sun_intensity  [3].
// The visibility buffer of the view is held fixed (computed by a previous
// frame).  Output: radiance [height, width, 4] (alpha constant 1).
#include <torch/extension.h>

#include <cstdint>

#include "lh/lh.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

const lh_scene* scene_of(int64_t a) { return reinterpret_cast<const lh_scene*>(static_cast<uintptr_t>(a)); }
const lh_view* view_of(int64_t a) { return reinterpret_cast<const lh_view*>(static_cast<uintptr_t>(a)); }
lh_stream stream_of(int64_t a) { return reinterpret_cast<lh_stream>(static_cast<uintptr_t>(a)); }

void check_param(const Tensor& t, int64_t n, const char* what) {
    TORCH_CHECK(t.is_cuda() && t.scalar_type() == torch::kFloat32 && t.is_contiguous(), what,
                " must be a contiguous float32 device tensor");
    TORCH_CHECK(t.numel() == n, what, " has ", t.numel(), " elements, expected ", n);
}

lh_resolve_params params_of(const Tensor& bc, const Tensor& r, const Tensor& m, const Tensor& sd, const Tensor& si,
                            const Tensor& frozen) {
    return lh_resolve_params{bc.data_ptr<float>(), r.data_ptr<float>(), m.data_ptr<float>(), sd.data_ptr<float>(),
                             si.data_ptr<float>(), reinterpret_cast<uint32_t*>(frozen.data_ptr<int32_t>())};
}

class ResolveFunction : public torch::autograd::Function<ResolveFunction> {
   public:
    static Tensor forward(AutogradContext* ctx, int64_t scene, int64_t view, int64_t stream, Tensor geometry,
                          int64_t geometry_base, Tensor base_color, Tensor roughness, Tensor metallic, Tensor sun_dir,
                          Tensor sun_intensity) {
        const lh_scene* s = scene_of(scene);
        const lh_view* v = view_of(view);
        TORCH_CHECK(s && v, "null scene or view");
        const int64_t K = s->num_mesh_infos;
        check_param(base_color, 3 * K, "base_color");
        check_param(roughness, K, "roughness");
        check_param(metallic, K, "metallic");
        check_param(sun_dir, 3, "sun_dir");
        check_param(sun_intensity, 3, "sun_intensity");
        TORCH_CHECK(geometry.is_cuda() && geometry.scalar_type() == torch::kFloat32 && geometry.is_contiguous(),
                    "geometry must be the scene's geometry arena as contiguous float32 words");
        TORCH_CHECK(uint64_t(geometry.data_ptr<float>()) == uint64_t(geometry_base),
                    "geometry must alias the scene's geometry arena");
        // per-pixel quantities the adjoint holds fixed, recorded by the forward
        Tensor frozen = torch::empty({int64_t(v->height) * v->width * 4},
                                     torch::TensorOptions().dtype(torch::kInt32).device(base_color.device()));
        const lh_resolve_params p = params_of(base_color, roughness, metallic, sun_dir, sun_intensity, frozen);
        const lh_status st = lh_resolve_ex(s, v, &p, stream_of(stream));
        TORCH_CHECK(st == LH_OK, "lh_resolve_ex failed with status ", st);
        auto opts = torch::TensorOptions().dtype(torch::kFloat32).device(base_color.device());
        Tensor out = torch::from_blob(v->radiance, {int64_t(v->height), int64_t(v->width), 4}, opts).clone();
        ctx->saved_data["scene"] = scene;
        ctx->saved_data["view"] = view;
        ctx->saved_data["stream"] = stream;
        ctx->saved_data["geometry_base"] = geometry_base;
        ctx->save_for_backward({geometry, base_color, roughness, metallic, sun_dir, sun_intensity, frozen});
        return out;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grad_outputs) {
        const lh_scene* s = scene_of(ctx->saved_data["scene"].toInt());
        const lh_view* v = view_of(ctx->saved_data["view"].toInt());
        const int64_t stream = ctx->saved_data["stream"].toInt();
        const uint64_t geometry_base = uint64_t(ctx->saved_data["geometry_base"].toInt());
        auto saved = ctx->get_saved_variables();
        const Tensor &geometry = saved[0], &bc = saved[1], &r = saved[2], &m = saved[3], &sd = saved[4],
                     &si = saved[5], &frozen = saved[6];
        Tensor g = grad_outputs[0].contiguous();
        TORCH_CHECK(g.numel() == int64_t(v->width) * v->height * 4, "radiance gradient of the wrong size");
        Tensor g_geo = torch::zeros_like(geometry), g_bc = torch::zeros_like(bc), g_r = torch::zeros_like(r),
               g_m = torch::zeros_like(m), g_sd = torch::zeros_like(sd), g_si = torch::zeros_like(si);
        const lh_resolve_params p = params_of(bc, r, m, sd, si, frozen);
        const lh_resolve_grads gr{g_bc.data_ptr<float>(), g_r.data_ptr<float>(), g_m.data_ptr<float>(),
                                  g_sd.data_ptr<float>(), g_si.data_ptr<float>(), g_geo.data_ptr<float>(),
                                  geometry_base, uint64_t(geometry.numel()) * 4};
        const lh_status st = lh_resolve_backward(s, v, &p, g.data_ptr<float>(), &gr, stream_of(stream));
        TORCH_CHECK(st == LH_OK, "lh_resolve_backward failed with status ", st);
        // order of forward's arguments: scene, view, stream, geometry, geometry_base, base_color, roughness,
        // metallic, sun_dir, sun_intensity
        return {Tensor(), Tensor(), Tensor(), g_geo, Tensor(), g_bc, g_r, g_m, g_sd, g_si};
    }
};

Tensor resolve(int64_t scene, int64_t view, int64_t stream, Tensor geometry, int64_t geometry_base, Tensor base_color,
               Tensor roughness, Tensor metallic, Tensor sun_dir, Tensor sun_intensity) {
    return ResolveFunction::apply(scene, view, stream, geometry, geometry_base, base_color, roughness, metallic,
                                  sun_dir, sun_intensity);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Differentiable visibility-buffer resolve (torch::autograd::Function over liblh.so)";
    m.def("resolve", &resolve, "radiance = resolve(scene*, view*, stream, geometry, geometry_base, base_color, "
                               "roughness, metallic, sun_dir, sun_intensity); differentiable in the tensors");
    m.def("version", []() { return std::string(lh_version()); });
}
