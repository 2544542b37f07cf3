// torch_render.cpp - libtorch autograd Function around vrt_render_fwd / vrt_render_bwd.
//
// The Python module _vrt_torch exposes render(albedo, world, cameras, ...),
// differentiable in `albedo` (fp32 [num_voxel_slots * 512, 3] on the device,
// voxel id = slot * 512 + voxel index in the brick).  Forward traces one ray
// per pixel and shades albedo * max(0, N.L) * intensity; backward sums dL/dpixel x weight
// per hit voxel into the albedo gradient: a dense gradient written whole (the default), a
// zero-filled one scatter-added with atomics, or a sparse COO gradient with one entry per
// pixel.  All device work goes through the C ABI of libvoxelrt_hip.so; torch only owns the
// tensors.
#include <torch/extension.h>

#include <cstring>
#include <stdexcept>
#include <vector>

#include "vrt.h"

namespace {

void check(int status, const char* what) {
    if (status != VRT_OK) throw std::runtime_error(std::string(what) + ": " + vrt_last_error());
}

std::vector<vrt_camera> unpack_cameras(const torch::Tensor& cams) {
    TORCH_CHECK(cams.device().is_cpu() && cams.scalar_type() == torch::kUInt8 && cams.is_contiguous(),
                "cameras: contiguous uint8 CPU tensor of packed vrt_camera structs");
    TORCH_CHECK(cams.numel() % sizeof(vrt_camera) == 0, "cameras: size is not a multiple of sizeof(vrt_camera)");
    std::vector<vrt_camera> out(cams.numel() / sizeof(vrt_camera));
    if (!out.empty()) std::memcpy(out.data(), cams.data_ptr(), cams.numel());
    return out;
}

// The kernels index albedo by slot * 512 + voxel for every allocated slot, so it must
// cover the world's slot high-water mark (the C ABI cannot check a raw pointer's length).
void check_inputs(const torch::Tensor& albedo, int64_t world, const std::vector<double>& light) {
    TORCH_CHECK(albedo.is_cuda() && albedo.scalar_type() == torch::kFloat32 && albedo.dim() == 2 &&
                    albedo.size(1) == 3,
                "albedo: float32 [voxels, 3] on the GPU");
    TORCH_CHECK(light.size() == 3, "light: 3 components");
    vrt_world_info info;
    check(vrt_world_get_info((const vrt_world*)world, &info), "vrt_world_get_info");
    TORCH_CHECK(albedo.size(0) >= (int64_t)info.slot_high_water * 512,
                "albedo: needs slot_high_water * 512 = ", (int64_t)info.slot_high_water * 512, " voxels, has ",
                albedo.size(0));
}

// How the backward pass forms the albedo gradient.
enum GradMode : int64_t {
    GRAD_BRICKS = 0,  // dense, written whole by vrt_render_bwd_bricks (the default)
    GRAD_ATOMIC = 1,  // dense, zero-filled then scatter-added by vrt_render_bwd
    GRAD_SPARSE = 2   // sparse COO, one unsummed entry per pixel (vrt_render_bwd_sparse)
};

class RenderFunction : public torch::autograd::Function<RenderFunction> {
public:
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor albedo, int64_t world,
                                 torch::Tensor cameras, int64_t max_iters, std::vector<double> light,
                                 double intensity, int64_t stream, int64_t grad_mode) {
        TORCH_CHECK(grad_mode >= GRAD_BRICKS && grad_mode <= GRAD_SPARSE, "grad_mode: 0 (bricks), 1 (atomic), 2 (sparse)");
        check_inputs(albedo, world, light);
        auto cams = unpack_cameras(cameras);
        int64_t npix = 0;
        for (const auto& c : cams) npix += (int64_t)c.width * c.height;
        auto alb = albedo.contiguous();
        auto image = torch::empty({npix, 3}, alb.options());
        auto saved_voxel = torch::empty({npix}, alb.options().dtype(torch::kInt32));
        auto saved_weight = torch::empty({npix}, alb.options());
        float l[3] = {(float)light[0], (float)light[1], (float)light[2]};
        check(vrt_render_fwd((const vrt_world*)world, cams.data(), (uint32_t)cams.size(), (uint32_t)max_iters, l,
                             (float)intensity, alb.data_ptr<float>(), image.data_ptr<float>(),
                             (uint32_t*)saved_voxel.data_ptr<int32_t>(), saved_weight.data_ptr<float>(),
                             (void*)stream),
              "vrt_render_fwd");
        ctx->save_for_backward({saved_voxel, saved_weight});
        ctx->saved_data["world"] = world;
        ctx->saved_data["stream"] = stream;
        ctx->saved_data["nvox"] = albedo.size(0);
        ctx->saved_data["mode"] = grad_mode;
        return image;
    }

    static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                   torch::autograd::variable_list grad_outputs) {
        auto saved = ctx->get_saved_variables();
        auto saved_voxel = saved[0], saved_weight = saved[1];
        int64_t world = ctx->saved_data["world"].toInt();
        int64_t stream = ctx->saved_data["stream"].toInt();
        int64_t nvox = ctx->saved_data["nvox"].toInt();
        int64_t mode = ctx->saved_data["mode"].toInt();
        auto g = grad_outputs[0].contiguous();
        const auto* w = (const vrt_world*)world;
        const uint32_t np = (uint32_t)saved_voxel.numel();
        const auto* sv = (const uint32_t*)saved_voxel.data_ptr<int32_t>();
        torch::Tensor grad;
        if (mode == GRAD_SPARSE) {
            auto index = torch::empty({1, (int64_t)np}, g.options().dtype(torch::kInt64));
            auto values = torch::empty({(int64_t)np, 3}, g.options());
            check(vrt_render_bwd_sparse(w, np, g.data_ptr<float>(), sv, saved_weight.data_ptr<float>(),
                                        index.data_ptr<int64_t>(), values.data_ptr<float>(), (void*)stream),
                  "vrt_render_bwd_sparse");
            // indices are in range by construction (saved ids < slot_high_water * 512 <= nvox):
            // no validation pass, which would read them back to the host
            grad = at::_sparse_coo_tensor_unsafe(index, values, {nvox, 3});
        } else if (mode == GRAD_ATOMIC) {
            grad = torch::zeros({nvox, 3}, g.options());
            check(vrt_render_bwd(w, np, g.data_ptr<float>(), sv, saved_weight.data_ptr<float>(), grad.data_ptr<float>(),
                                 (void*)stream),
                  "vrt_render_bwd");
        } else {
            grad = torch::empty({nvox, 3}, g.options());
            size_t bytes = vrt_render_bwd_workspace_bytes(np, (uint64_t)nvox);
            auto ws = torch::empty({(int64_t)bytes}, g.options().dtype(torch::kUInt8));
            check(vrt_render_bwd_bricks(w, np, g.data_ptr<float>(), sv, saved_weight.data_ptr<float>(), (uint64_t)nvox,
                                        grad.data_ptr<float>(), ws.data_ptr(), bytes, (void*)stream),
                  "vrt_render_bwd_bricks");
        }
        return {grad, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor render(torch::Tensor albedo, int64_t world, torch::Tensor cameras, int64_t max_iters,
                     std::vector<double> light, double intensity, int64_t stream, int64_t grad_mode) {
    return RenderFunction::apply(albedo, world, cameras, max_iters, light, intensity, stream, grad_mode);
}

// Forward only, also returning the saved hit voxel ids and shading weights.
std::vector<torch::Tensor> render_saved(torch::Tensor albedo, int64_t world, torch::Tensor cameras, int64_t max_iters,
                                        std::vector<double> light, double intensity, int64_t stream) {
    check_inputs(albedo, world, light);
    auto cams = unpack_cameras(cameras);
    int64_t npix = 0;
    for (const auto& c : cams) npix += (int64_t)c.width * c.height;
    auto alb = albedo.contiguous();
    auto image = torch::empty({npix, 3}, alb.options());
    auto sv = torch::empty({npix}, alb.options().dtype(torch::kInt32));
    auto sw = torch::empty({npix}, alb.options());
    float l[3] = {(float)light[0], (float)light[1], (float)light[2]};
    check(vrt_render_fwd((const vrt_world*)world, cams.data(), (uint32_t)cams.size(), (uint32_t)max_iters, l,
                         (float)intensity, alb.data_ptr<float>(), image.data_ptr<float>(),
                         (uint32_t*)sv.data_ptr<int32_t>(), sw.data_ptr<float>(), (void*)stream),
          "vrt_render_fwd");
    return {image, sv, sw};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("render", &render, "differentiable Lambertian render (libtorch autograd Function)");
    m.def("render_saved", &render_saved, "forward render returning image, saved voxel ids and weights");
}
