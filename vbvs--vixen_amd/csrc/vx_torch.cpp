// vx_torch.cpp -- libtorch binding of the differentiable albedo renderer:
// a torch::autograd::Function whose forward is vx_albedo_forward and whose
// backward is the hand-written adjoint vx_albedo_backward (both in
// libvixen_hip.so, called through the C ABI on the caller's stream).
//
//   image = albedo_render(albedo [V,3] f32 cuda, scene handle, camera (16 f32, cpu),
//                         width, height, light (3 f32, cpu), store_hits, deterministic, stream)
//
// store_hits keeps each pixel's albedo row and shade (8 bytes per pixel) for
// the backward pass; without it the backward re-traces the hits.
// deterministic selects the sorted segmented reduction over FP32 atomics.
// The discrete hit selection is not differentiated: only albedo gets a gradient.
//
// Streams. The tensors made here (image, the stored hits, the gradient) come
// from torch's caching allocator on the current stream, so the kernels must
// run on that stream: forward requires `stream` to be the current stream, and
// backward runs on the current stream of the backward pass (the engine makes
// it the stream the forward ran on), which is also where torch expects the
// gradient to be produced.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <stdexcept>
#include <string>

#include "vixen_hip.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(vx_status st, const char* where)
{
    if (st != VX_OK) throw std::runtime_error(std::string(where) + ": " + vx_last_error());
}

vx_camera to_camera(const torch::Tensor& cam)
{
    TORCH_CHECK(cam.device().is_cpu() && cam.scalar_type() == torch::kFloat32 && cam.numel() == 16,
                "camera must be 16 float32 values on the CPU (the 64-byte push constant)");
    vx_camera c;
    std::memcpy(&c, cam.contiguous().data_ptr<float>(), sizeof c);
    return c;
}

// A one-byte CPU tensor whose storage holds a reference to the scene. Kept in
// the autograd node, it keeps the scene usable until the node is released,
// whether or not its backward ever runs, even if the caller destroys the scene
// first.
torch::Tensor scene_keepalive(vx_scene* sc)
{
    check(vx_scene_retain(sc), "vx_scene_retain");
    auto* token = new uint8_t(0);
    return torch::from_blob(
        token, {1},
        [sc](void* p) {
            delete static_cast<uint8_t*>(p);
            (void)vx_scene_destroy(sc);
        },
        torch::TensorOptions().dtype(torch::kUInt8));
}

struct AlbedoRender : public torch::autograd::Function<AlbedoRender> {
    static torch::Tensor forward(AutogradContext* ctx, torch::Tensor albedo, int64_t scene, torch::Tensor camera,
                                 int64_t width, int64_t height, torch::Tensor light, bool store_hits,
                                 bool deterministic, int64_t stream)
    {
        TORCH_CHECK(albedo.is_cuda() && albedo.scalar_type() == torch::kFloat32 && albedo.dim() == 2 &&
                        albedo.size(1) == 3,
                    "albedo must be a [V,3] float32 device tensor");
        TORCH_CHECK(light.device().is_cpu() && light.numel() == 3 && light.scalar_type() == torch::kFloat32,
                    "light must be 3 float32 values on the CPU");
        const hipStream_t cur = c10::hip::getCurrentHIPStream(albedo.device().index()).stream();
        TORCH_CHECK(reinterpret_cast<hipStream_t>(stream) == cur,
                    "albedo_render: stream must be the current stream of the albedo's device (enter "
                    "torch.cuda.stream(s) around the call to use another)");
        auto* sc = reinterpret_cast<vx_scene*>(scene);
        uint64_t voxels = 0;
        check(vx_scene_voxel_count(sc, &voxels), "vx_scene_voxel_count");
        TORCH_CHECK(static_cast<uint64_t>(albedo.size(0)) == voxels, "albedo has ", albedo.size(0),
                    " rows, the scene has ", voxels, " occupied voxels");
        const vx_camera cam = to_camera(camera);
        auto alb = albedo.contiguous();
        auto lt = light.contiguous();
        auto image = torch::empty({height, width, 3}, alb.options());
        torch::Tensor rows, shade;
        if (store_hits) {
            rows = torch::empty({height * width}, alb.options().dtype(torch::kInt32));
            shade = torch::empty({height * width}, alb.options());
        }
        check(vx_albedo_forward(sc, &cam, static_cast<uint32_t>(width), static_cast<uint32_t>(height),
                                lt.data_ptr<float>(), alb.data_ptr<float>(), image.data_ptr<float>(),
                                store_hits ? reinterpret_cast<uint32_t*>(rows.data_ptr<int32_t>()) : nullptr,
                                store_hits ? shade.data_ptr<float>() : nullptr, reinterpret_cast<vx_stream>(stream)),
              "vx_albedo_forward");
        ctx->save_for_backward({rows, shade, camera, lt});
        ctx->saved_data["scene"] = scene;
        ctx->saved_data["scene_ref"] = scene_keepalive(sc);
        ctx->saved_data["width"] = width;
        ctx->saved_data["height"] = height;
        ctx->saved_data["voxels"] = static_cast<int64_t>(voxels);
        ctx->saved_data["deterministic"] = deterministic;
        return image;
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grads)
    {
        auto saved = ctx->get_saved_variables();
        const auto& rows = saved[0];
        const auto& shade = saved[1];
        const vx_camera cam = to_camera(saved[2]);
        const auto& lt = saved[3];
        auto* sc = reinterpret_cast<vx_scene*>(ctx->saved_data["scene"].toInt());
        const int64_t w = ctx->saved_data["width"].toInt(), h = ctx->saved_data["height"].toInt();
        const int64_t voxels = ctx->saved_data["voxels"].toInt();
        const bool det = ctx->saved_data["deterministic"].toBool();
        auto g = grads[0].contiguous();
        const hipStream_t stream = c10::hip::getCurrentHIPStream(g.device().index()).stream();
        auto grad = torch::zeros({voxels, 3}, g.options());
        const bool stored = rows.defined();
        check(vx_albedo_backward(sc, &cam, static_cast<uint32_t>(w), static_cast<uint32_t>(h), lt.data_ptr<float>(),
                                 g.data_ptr<float>(),
                                 stored ? reinterpret_cast<const uint32_t*>(rows.data_ptr<int32_t>()) : nullptr,
                                 stored ? shade.data_ptr<float>() : nullptr, grad.data_ptr<float>(),
                                 det ? VX_GRAD_DETERMINISTIC : VX_GRAD_ATOMIC, reinterpret_cast<vx_stream>(stream)),
              "vx_albedo_backward");
        return {grad, torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor(),
                torch::Tensor(), torch::Tensor(), torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor albedo_render(torch::Tensor albedo, int64_t scene, torch::Tensor camera, int64_t width, int64_t height,
                            torch::Tensor light, bool store_hits, bool deterministic, int64_t stream)
{
    return AlbedoRender::apply(albedo, scene, camera, width, height, light, store_hits, deterministic, stream);
}

} // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "libtorch autograd binding of the HIP albedo renderer (forward + hand-written adjoint)";
    m.def("albedo_render", &albedo_render, "Differentiable Lambertian render of per-voxel albedo");
}
