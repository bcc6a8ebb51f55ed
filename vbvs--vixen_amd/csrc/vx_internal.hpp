// vx_internal.hpp -- library-private definitions shared by the HIP translation
// units of libvixen_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "vixen_hip.h"
#include "vx_common.hpp"
#include "vx_trace.hpp"

struct vx_albedo_state;
void vx_albedo_state_free(vx_albedo_state* a);
namespace vx {
// Looks up the device address of the no-allocation workspace on `device` (current).
void scratch_init(int device);
} // namespace vx

struct vx_ctx {
    int device;
    int cu_count;
};

// Scenes and batches copy what they need of their context (its device and CU
// count), so a context may be destroyed before them.
struct vx_scene {
    int device;
    int cu_count;
    uint2* nodes;
    uint32_t* brick_data;
    uint32_t* occ;
    uint32_t* box; // per brick: box of its occupied voxels (see SceneView::box)
    uint2* bc1;
    uint4* nrm; // DXT normal blocks (32 per brick) or nullptr
    float* palette = nullptr; // 256 x RGB material palette (always the scene's own copy) or nullptr
    uint64_t node_count;
    uint64_t brick_count;
    vx_octree_config cfg;
    vx::SceneView view;
    // descriptors in the first k+1 breadth-first levels, k = 0..3 (a prefix of
    // the descriptor array when it is in breadth-first order)
    uint32_t level_prefix[4];
    uint32_t lds_depth; // levels staged in LDS by vx_trace / vx_render (0 = none)
    vx_albedo_state* albedo = nullptr; // differentiable-rendering state, built on first use
    // vx_scene_create_device: nodes, brick_data, bc1 and nrm are the caller's
    // device buffers (not freed here); occ and box are always the scene's own
    bool borrowed = false;
    // references: the creator's, plus one per vx_scene_retain; vx_scene_destroy
    // drops one and the last frees the scene
    std::atomic<int> refs{1};
    // set when a vx_scene_revalidate finds the buffers broken: launches refuse the scene
    std::atomic<bool> invalid{false};
};

namespace vx {

#define VX_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return ::vx::fail(VX_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

// Makes a context's device current for one API call and gives the caller's
// current device back on return: null-stream launches, allocations and copies
// then go to the device the scene lives on, whatever device the calling thread
// had selected.
class DeviceGuard {
public:
    explicit DeviceGuard(int device)
    {
        int cur = -1;
        err_ = hipGetDevice(&cur);
        if (err_ == hipSuccess && cur != device) {
            err_ = hipSetDevice(device);
            if (err_ == hipSuccess) prev_ = cur;
        }
    }
    ~DeviceGuard()
    {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }
    DeviceGuard(const DeviceGuard&) = delete;
    DeviceGuard& operator=(const DeviceGuard&) = delete;
    hipError_t error() const { return err_; }

private:
    int prev_ = -1;
    hipError_t err_ = hipSuccess;
};

#define VX_DEVICE_GUARD(device)                                                                        \
    ::vx::DeviceGuard vx_device_guard_(device);                                                       \
    if (vx_device_guard_.error() != hipSuccess)                                                       \
        return ::vx::fail(VX_ERR_HIP, std::string("could not make device ") + std::to_string(device) + \
                                          " current: " + hipGetErrorString(vx_device_guard_.error()))

// VX_OK, or the error every launch on a scene that failed vx_scene_revalidate returns.
inline vx_status scene_usable(const vx_scene* s, const char* who)
{
    if (!s->invalid.load(std::memory_order_acquire)) return VX_OK;
    return fail(VX_ERR_INVALID_SCENE, std::string(who) + ": the scene failed its last vx_scene_revalidate");
}

constexpr uint32_t kXcds = 8;
constexpr uint32_t kMaxLdsBytes = 65536;

// Chunk of workgroup blockIdx.x when the XCDs take stripes of S consecutive
// chunks in turn: XCD x runs workgroups i = 8k + x, and its k-th takes chunk
// k % S of stripe 8 (k / S) + x; chunks past the last whole round of stripes go
// to the XCDs in bands. Neighbouring rays still share an XCD's L2, while every
// part of the image or ray set -- sky rows and ground rows, a call's camera
// rays and its incoherent rays -- is spread over every XCD (contiguous bands
// leave XCDs that drew cheap rays idle while the others finish).
template <uint32_t S>
__device__ __forceinline__ uint32_t xcd_stripe_chunk(uint32_t per_xcd)
{
    const uint32_t x = blockIdx.x % kXcds, k = blockIdx.x / kXcds;
    const uint32_t full = per_xcd / S * S;
    if (k < full) return ((k / S) * kXcds + x) * S + k % S;
    return full * kXcds + x * (per_xcd - full) + (k - full);
}

// Stripe length of the image kernels (vx_render, the albedo passes): 16x16 pixel tiles.
constexpr uint32_t kRenderStripe = 16;

// Staged descriptor count for `depth` levels of `s` (0 = none).
inline uint32_t staged_nodes(const vx_scene* s, uint32_t depth)
{
    if (depth == 0) return 0;
    const uint32_t d = depth > 4 ? 4 : depth;
    return s->level_prefix[d - 1];
}

} // namespace vx
