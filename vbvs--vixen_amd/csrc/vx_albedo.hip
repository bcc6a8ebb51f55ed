// vx_albedo.hip -- differentiable Lambertian albedo rendering and its adjoint.
//
// No reference counterpart: VIXEN renders only. The forward pass reuses the
// traversal of vx_render unchanged (the discrete hit selection is not
// differentiated); per pixel it shades a per-voxel FP32 RGB albedo,
//   pixel = albedo[v] * max(dot(n, light), 0),
// with n the entry-face normal of the hit voxel (the axis the DDA crossed
// into it) and light a fixed unit vector; misses are 0. Parameters are one
// row per occupied voxel: row(v) = base[brick] + (occupied voxels before v in
// the brick's linear order), read from the occupancy mask.
//
// The adjoint pushes dL/dpixel through the shading into dL/dalbedo:
//   dL/dalbedo[v] += dL/dpixel[p] * shade[p]   for every pixel p that hit v.
// Hits are either read back from the forward pass (8 bytes per pixel) or
// recomputed by re-tracing (no per-pixel storage). Two reductions:
//   atomic         one FP32 atomic add per hit pixel and channel;
//   deterministic  (row, pixel) pairs are radix-sorted by row (stable, so
//                  equal rows keep pixel order, misses last); the run heads
//                  are selected (integer work) and each run is summed in a
//                  fixed order (one thread in sequence for runs of <= 32 pairs,
//                  else one workgroup: thread t takes elements t, t+256, ...,
//                  then a fixed tree), so the per-row sums, added without
//                  conflicts, depend only on the input: identical bits on
//                  every run. (A decoupled look-back reduce-by-key is not
//                  used: the order in which it combines the partial sums of a
//                  run spanning several tiles depends on timing.)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>

#include "vx_internal.hpp"

namespace vx {

constexpr uint32_t kAlbBlock = 256;

// LSD radix sort (onesweep) at every size: rocprim's default switches to a
// block-sort + merge-pass algorithm below 2^20 items, which at one 1280x720
// view costs about 20 launches instead of 3 onesweep passes over the key bits.
using PairSortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                  rocprim::default_config, 0>;
constexpr uint32_t kNoVoxel = 0xFFFFFFFFu;

struct Rgb {
    float r, g, b;
};

struct AlbCam {
    float pos[3], dir[3], up[3], right[3];
    float tan_half, aspect;
    float light[3];
};

__global__ void __launch_bounds__(kAlbBlock) k_brick_popc(const uint32_t* __restrict__ occ, uint64_t nb,
                                                         uint32_t* __restrict__ counts)
{
    const uint64_t b = (uint64_t)blockIdx.x * kAlbBlock + threadIdx.x;
    if (b >= nb) return;
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) c += __builtin_popcount(occ[b * 16 + w]);
    counts[b] = c;
}

// Row of the hit voxel: brick base plus occupied voxels before `lin` in the brick.
__device__ __forceinline__ uint32_t voxel_row(const SceneView& s, const uint32_t* __restrict__ base, uint32_t brick,
                                              uint32_t lin)
{
    const uint32_t* m = s.occ + (size_t)brick * 16u;
    const uint32_t wq = lin >> 5;
    uint32_t r = base[brick];
    for (uint32_t w = 0; w < wq; ++w) r += __builtin_popcount(m[w]);
    return r + __builtin_popcount(m[wq] & ((1u << (lin & 31u)) - 1u));
}

__device__ __forceinline__ float face_shade(uint32_t flags, const float light[3])
{
    const uint32_t axis = (flags >> 1) & 3u;
    const float l0 = light[0], l1 = light[1], l2 = light[2];
    const float l = axis == 0u ? l0 : (axis == 1u ? l1 : l2);
    const float ndl = (flags & 8u) ? l : -l;
    return ndl > 0.0f ? ndl : 0.0f;
}

// Traces pixel p of a w x h image; returns its albedo row (kNoVoxel on a miss) and shade.
__device__ __forceinline__ uint32_t trace_pixel(const SceneView& s, const AlbCam& c, const uint32_t* base,
                                                uint32_t px, uint32_t py, uint32_t w, uint32_t h, float& shade)
{
    const float u = ((float)px + 0.5f) / (float)w;
    const float v = ((float)py + 0.5f) / (float)h;
    const float ndcx = u * 2.0f - 1.0f;
    const float ndcy = -(v * 2.0f - 1.0f);
    float d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a = ((c.right[i] * ndcx) * c.tan_half) * c.aspect;
        const float b = (c.up[i] * ndcy) * c.tan_half;
        d[i] = (c.dir[i] + a) + b;
    }
    vx_hit hr;
    TraceCounts cnt{0u, 0u, 0u, 0u};
    shade = 0.0f;
    if (!trace_ray<false, false, false>(s, c.pos, d, 0.0f, 3.0e38f, hr, cnt)) return kNoVoxel;
    shade = face_shade(hr.flags, c.light);
    return voxel_row(s, base, hr.brick, hr.voxel);
}

// Pixel of lane threadIdx.x in 16x16 tile blockIdx.x (row-major tiles, 8x8 per wave).
__device__ __forceinline__ bool tile_pixel(uint32_t w, uint32_t h, uint32_t& px, uint32_t& py)
{
    const uint32_t tiles_x = (w + 15u) / 16u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    px = (blockIdx.x % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    py = (blockIdx.x / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
    return px < w && py < h;
}

__global__ void __launch_bounds__(kAlbBlock)
    k_albedo_forward(SceneView s, AlbCam c, uint32_t w, uint32_t h, const uint32_t* __restrict__ base,
                     const float* __restrict__ albedo, float* __restrict__ image, uint32_t* __restrict__ hit_row,
                     float* __restrict__ hit_shade)
{
    uint32_t px, py;
    if (!tile_pixel(w, h, px, py)) return;
    float shade;
    const uint32_t row = trace_pixel(s, c, base, px, py, w, h, shade);
    const uint64_t p = (uint64_t)py * w + px;
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (row != kNoVoxel) {
        r = albedo[(size_t)row * 3 + 0] * shade;
        g = albedo[(size_t)row * 3 + 1] * shade;
        b = albedo[(size_t)row * 3 + 2] * shade;
    }
    image[p * 3 + 0] = r;
    image[p * 3 + 1] = g;
    image[p * 3 + 2] = b;
    if (hit_row) {
        hit_row[p] = row;
        hit_shade[p] = shade;
    }
}

// Sum of v over the lanes in `group` (a subset of the active lanes), broadcast to all lanes.
__device__ __forceinline__ float group_sum(float v, uint64_t group, uint32_t lane)
{
    float x = ((group >> lane) & 1ull) ? v : 0.0f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Per-pixel contribution dL/dpixel * shade. ATOMIC adds it into grad with wave
// aggregation: lanes that hit the same row are summed across the wave first, so
// a voxel near the camera covering thousands of pixels receives one hardware
// FP32 atomic per wave instead of one per pixel (fewer contended atomics and
// less rounding in the atomic unit); once the next group is a single lane the
// remaining lanes add directly. Otherwise it emits (row, contribution) pairs
// for the sorted reduction. Every lane runs to the end (no early exit), so the
// cross-lane operations see the whole wave.
template <bool RECOMPUTE, bool ATOMIC>
__global__ void __launch_bounds__(kAlbBlock)
    k_albedo_adjoint(SceneView s, AlbCam c, uint32_t w, uint32_t h, const uint32_t* __restrict__ base,
                     const float* __restrict__ grad_image, const uint32_t* __restrict__ hit_row,
                     const float* __restrict__ hit_shade, float* __restrict__ grad, uint32_t* __restrict__ keys,
                     Rgb* __restrict__ vals, uint32_t* __restrict__ pix, uint32_t miss_key)
{
    uint32_t px, py;
    const bool inside = tile_pixel(w, h, px, py);
    const uint64_t p = (uint64_t)py * w + px;
    uint32_t row = kNoVoxel;
    float shade = 0.0f;
    if (inside) {
        if (RECOMPUTE) {
            row = trace_pixel(s, c, base, px, py, w, h, shade);
        } else {
            row = hit_row[p];
            shade = hit_shade[p];
        }
    }
    float gr = 0.0f, gg = 0.0f, gb = 0.0f;
    if (inside) {
        gr = grad_image[p * 3 + 0] * shade;
        gg = grad_image[p * 3 + 1] * shade;
        gb = grad_image[p * 3 + 2] * shade;
    }
    if (ATOMIC) {
        const uint32_t lane = threadIdx.x & 63u;
        uint64_t pending = __ballot(row != kNoVoxel);
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const uint32_t lrow = __shfl(row, leader, 64);
            const uint64_t group = __ballot(row == lrow) & pending;
            if ((group & (group - 1)) == 0ull) break; // a single lane: stop aggregating
            const float sr = group_sum(gr, group, lane), sg = group_sum(gg, group, lane),
                        sb = group_sum(gb, group, lane);
            if ((int)lane == leader) {
                atomicAdd(&grad[(size_t)lrow * 3 + 0], sr);
                atomicAdd(&grad[(size_t)lrow * 3 + 1], sg);
                atomicAdd(&grad[(size_t)lrow * 3 + 2], sb);
            }
            pending &= ~group;
        }
        if ((pending >> lane) & 1ull) {
            atomicAdd(&grad[(size_t)row * 3 + 0], gr);
            atomicAdd(&grad[(size_t)row * 3 + 1], gg);
            atomicAdd(&grad[(size_t)row * 3 + 2], gb);
        }
    } else if (inside) {
        keys[p] = row == kNoVoxel ? miss_key : row; // misses sort last
        vals[p] = Rgb{gr, gg, gb};
        pix[p] = (uint32_t)p;
    }
}

// Run heads of the sorted keys: index i starts a run when its key differs from
// the previous one.
struct RunHead {
    const uint32_t* keys;
    __host__ __device__ bool operator()(uint32_t i) const { return i == 0u || keys[i] != keys[i - 1u]; }
};

// Fixed-order run sums. Every run's sum is a function of its elements and
// length only, so the gradient bits do not depend on scheduling. Runs of at
// most kShortRun pairs: one thread adds them in order. Longer runs (voxels that
// cover many pixels) are listed for k_reduce_long_runs. The last run, of
// misses (key miss_key), is skipped. Rows are unique across runs: no conflicts.
constexpr uint32_t kShortRun = 32;

__device__ __forceinline__ uint32_t run_end(const uint32_t* heads, uint32_t r, uint32_t n, uint32_t n_items)
{
    return r + 1u < n ? heads[r + 1u] : n_items;
}

__global__ void __launch_bounds__(kAlbBlock) k_reduce_short_runs(const uint32_t* __restrict__ keys_sorted,
                                                                const Rgb* __restrict__ vals,
                                                                const uint32_t* __restrict__ pix_sorted,
                                                                const uint32_t* __restrict__ heads,
                                                                uint32_t* __restrict__ counts, uint32_t n_items,
                                                                uint32_t miss_key, float* __restrict__ grad,
                                                                uint32_t* __restrict__ long_runs)
{
    const uint32_t n = counts[0];
    for (uint32_t r = blockIdx.x * kAlbBlock + threadIdx.x; r < n; r += gridDim.x * kAlbBlock) {
        const uint32_t b = heads[r], e = run_end(heads, r, n, n_items);
        const uint32_t row = keys_sorted[b];
        if (row >= miss_key) continue;
        if (e - b > kShortRun) {
            long_runs[atomicAdd(&counts[1], 1u)] = r; // list order is irrelevant to the sums
            continue;
        }
        float sr = 0.0f, sg = 0.0f, sb = 0.0f;
        for (uint32_t i = b; i < e; ++i) {
            const Rgb v = vals[pix_sorted[i]];
            sr += v.r;
            sg += v.g;
            sb += v.b;
        }
        grad[(size_t)row * 3 + 0] += sr;
        grad[(size_t)row * 3 + 1] += sg;
        grad[(size_t)row * 3 + 2] += sb;
    }
}

// One workgroup per listed long run: thread t adds elements t, t+256, t+512,
// ... in order into two interleaved partial sums (independent loads in flight),
// then the threads' sums are combined by a fixed tree in LDS.
__global__ void __launch_bounds__(kAlbBlock) k_reduce_long_runs(const uint32_t* __restrict__ keys_sorted,
                                                               const Rgb* __restrict__ vals,
                                                               const uint32_t* __restrict__ pix_sorted,
                                                               const uint32_t* __restrict__ heads,
                                                               const uint32_t* __restrict__ counts, uint32_t n_items,
                                                               const uint32_t* __restrict__ long_runs,
                                                               float* __restrict__ grad)
{
    __shared__ float part[3][kAlbBlock];
    const uint32_t t = threadIdx.x;
    const uint32_t n = counts[0], n_long = counts[1];
    for (uint32_t j = blockIdx.x; j < n_long; j += gridDim.x) {
        const uint32_t r = long_runs[j]; // the same j, so the same run, in the whole group
        const uint32_t b = heads[r], e = run_end(heads, r, n, n_items);
        float s0[3] = {0.0f, 0.0f, 0.0f}, s1[3] = {0.0f, 0.0f, 0.0f};
        uint32_t i = b + t;
        for (; i + kAlbBlock < e; i += 2u * kAlbBlock) {
            const Rgb v0 = vals[pix_sorted[i]], v1 = vals[pix_sorted[i + kAlbBlock]];
            s0[0] += v0.r; s0[1] += v0.g; s0[2] += v0.b;
            s1[0] += v1.r; s1[1] += v1.g; s1[2] += v1.b;
        }
        if (i < e) {
            const Rgb v0 = vals[pix_sorted[i]];
            s0[0] += v0.r; s0[1] += v0.g; s0[2] += v0.b;
        }
        part[0][t] = s0[0] + s1[0];
        part[1][t] = s0[1] + s1[1];
        part[2][t] = s0[2] + s1[2];
        __syncthreads();
        for (uint32_t h = kAlbBlock / 2u; h > 0u; h >>= 1) {
            if (t < h) {
                part[0][t] += part[0][t + h];
                part[1][t] += part[1][t + h];
                part[2][t] += part[2][t + h];
            }
            __syncthreads();
        }
        if (t == 0u) {
            const uint32_t row = keys_sorted[b];
            grad[(size_t)row * 3 + 0] += part[0][0];
            grad[(size_t)row * 3 + 1] += part[1][0];
            grad[(size_t)row * 3 + 2] += part[2][0];
        }
        __syncthreads(); // part[] is reused by the next run
    }
}

__global__ void k_clear_long_count(uint32_t* counts) { counts[1] = 0u; }

// Device buffers of the deterministic reduction for up to cap_px pixels.
struct DetWorkspace {
    uint32_t *keys, *keys_sorted, *heads;
    uint32_t* n_runs; // [0] runs, [1] long runs
    Rgb* vals;        // contributions, by pixel
    uint32_t *pix, *pix_sorted;
    void* tmp;        // sort and run-selection temporary storage
    size_t cap_px, cap_tmp;
};

// Workspace of vx_albedo_backward_noalloc: one array per device, kept until the
// process ends, shared by every call on that device (calls that overlap on one
// device race on it). Host code never uses its name as an address: its device
// address comes from hipGetSymbolAddress, once per device, at vx_ctx_create.
constexpr size_t kScratchBytes = size_t{48} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_err[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};

} // namespace vx

// Per-scene state of the albedo path (built on first use).
struct vx_albedo_state {
    uint32_t* base = nullptr; // per-brick row base
    uint64_t voxels = 0;
    // sorted-reduction workspace, grown on demand
    vx::DetWorkspace ws{};
};

namespace vx {

static vx_albedo_state* albedo_state(vx_scene* s, hipStream_t st, vx_status& status)
{
    status = VX_OK;
    if (s->albedo) return s->albedo;
    auto* a = new (std::nothrow) vx_albedo_state{};
    if (!a) {
        status = fail(VX_ERR_OUT_OF_MEMORY, "albedo state: out of memory");
        return nullptr;
    }
    const uint64_t nb = s->brick_count;
    uint32_t* counts = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint32_t last[2] = {0, 0};
    bool ok = hipMalloc(&a->base, std::max<uint64_t>(1, nb) * 4) == hipSuccess &&
              hipMalloc(&counts, std::max<uint64_t>(1, nb) * 4) == hipSuccess;
    if (ok && nb) {
        k_brick_popc<<<dim3((unsigned)((nb + kAlbBlock - 1) / kAlbBlock)), dim3(kAlbBlock), 0, st>>>(s->occ, nb, counts);
        ok = hipGetLastError() == hipSuccess &&
             hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, a->base, (int)nb, st) == hipSuccess &&
             hipMalloc(&tmp, std::max<size_t>(1, tmp_bytes)) == hipSuccess &&
             hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, a->base, (int)nb, st) == hipSuccess &&
             hipMemcpyAsync(&last[0], a->base + nb - 1, 4, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipMemcpyAsync(&last[1], counts + nb - 1, 4, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    (void)hipFree(counts);
    (void)hipFree(tmp);
    if (!ok) {
        (void)hipFree(a->base);
        delete a;
        status = fail(VX_ERR_HIP, "albedo state: voxel row table failed");
        return nullptr;
    }
    a->voxels = nb ? (uint64_t)last[0] + last[1] : 0;
    s->albedo = a;
    return a;
}

static AlbCam make_alb_cam(const vx_camera* cam, const float light[3])
{
    AlbCam c;
    for (int i = 0; i < 3; ++i) {
        c.pos[i] = cam->pos[i];
        c.dir[i] = cam->dir[i];
        c.up[i] = cam->up[i];
        c.right[i] = cam->right[i];
        c.light[i] = light[i];
    }
    c.tan_half = std::tan((cam->fov_deg * 0.5f) * 0.017453292519943295f);
    c.aspect = cam->aspect;
    return c;
}

// The albedo kernels read descriptors from global memory (no staged prefix);
// their LDS holds the traversal stacks only.
static SceneView alb_view(const vx_scene* s, uint32_t& lds)
{
    SceneView v = s->view;
    v.lds_off = 0;
    v.lds_count = 0;
    lds = lay_out_lanes(v, 0, kAlbBlock, false);
    return v;
}

static vx_status grow(vx_albedo_state* a, size_t px, size_t tmp)
{
    DetWorkspace& w = a->ws;
    if (px > w.cap_px) {
        for (void* p : {(void*)w.keys, (void*)w.keys_sorted, (void*)w.heads, (void*)w.n_runs, (void*)w.vals,
                        (void*)w.pix, (void*)w.pix_sorted})
            (void)hipFree(p);
        w.keys = w.keys_sorted = w.heads = w.n_runs = w.pix = w.pix_sorted = nullptr;
        w.vals = nullptr;
        w.cap_px = 0;
        if (hipMalloc(&w.keys, px * 4) != hipSuccess || hipMalloc(&w.keys_sorted, px * 4) != hipSuccess ||
            hipMalloc(&w.heads, px * 4) != hipSuccess || hipMalloc(&w.n_runs, 8) != hipSuccess ||
            hipMalloc(&w.vals, px * sizeof(Rgb)) != hipSuccess || hipMalloc(&w.pix, px * 4) != hipSuccess ||
            hipMalloc(&w.pix_sorted, px * 4) != hipSuccess)
            return fail(VX_ERR_OUT_OF_MEMORY, "albedo workspace allocation failed");
        w.cap_px = px;
    }
    if (tmp > w.cap_tmp) {
        (void)hipFree(w.tmp);
        w.tmp = nullptr;
        w.cap_tmp = 0;
        if (hipMalloc(&w.tmp, tmp) != hipSuccess) return fail(VX_ERR_OUT_OF_MEMORY, "albedo workspace allocation failed");
        w.cap_tmp = tmp;
    }
    return VX_OK;
}

// The deterministic reduction's workspace for px pixels carved from the
// device's scratch array: 256-byte aligned slices, the temporary storage last.
// Fails when px pixels and tmp bytes do not fit.
static vx_status scratch_workspace(int device, size_t px, size_t tmp, DetWorkspace& w)
{
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device < 0 || device >= kMaxDevices || !g_scratch_done[device])
            return fail(VX_ERR_INVALID_ARG, "vx_albedo_backward_noalloc: no context was created on device " +
                                                std::to_string(device));
        if (g_scratch_err[device] != hipSuccess)
            return fail(VX_ERR_HIP, std::string("vx_albedo_backward_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                        hipGetErrorString(g_scratch_err[device]));
        base = g_scratch_ptr[device];
    }
    auto* p = static_cast<unsigned char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* q = p + off;
        off += (bytes + 255) & ~size_t{255};
        return q;
    };
    w.keys = static_cast<uint32_t*>(take(px * 4));
    w.keys_sorted = static_cast<uint32_t*>(take(px * 4));
    w.heads = static_cast<uint32_t*>(take(px * 4));
    w.n_runs = static_cast<uint32_t*>(take(8));
    w.vals = static_cast<Rgb*>(take(px * sizeof(Rgb)));
    w.pix = static_cast<uint32_t*>(take(px * 4));
    w.pix_sorted = static_cast<uint32_t*>(take(px * 4));
    w.tmp = take(tmp);
    w.cap_px = px;
    w.cap_tmp = tmp;
    if (off > kScratchBytes)
        return fail(VX_ERR_INVALID_ARG, "vx_albedo_backward_noalloc: " + std::to_string(px) + " pixels need " +
                                            std::to_string(off) + " bytes of workspace, the scratch array holds " +
                                            std::to_string(kScratchBytes));
    return VX_OK;
}

} // namespace vx

void vx_albedo_state_free(vx_albedo_state* a)
{
    if (!a) return;
    const vx::DetWorkspace& w = a->ws;
    for (void* p : {(void*)a->base, (void*)w.keys, (void*)w.keys_sorted, (void*)w.heads, (void*)w.n_runs,
                    (void*)w.vals, (void*)w.pix, (void*)w.pix_sorted, w.tmp})
        (void)hipFree(p);
    delete a;
}

extern "C" {

vx_status vx_scene_voxel_count(vx_scene* s, uint64_t* n_voxels)
{
    if (!s || !n_voxels) return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_voxel_count: null argument");
    VX_DEVICE_GUARD(s->device);
    if (vx_status us = vx::scene_usable(s, "vx_albedo"); us != VX_OK) return us;
    vx_status st;
    vx_albedo_state* a = vx::albedo_state(s, nullptr, st);
    if (!a) return st;
    *n_voxels = a->voxels;
    return VX_OK;
}

vx_status vx_albedo_forward(vx_scene* s, const vx_camera* cam, uint32_t w, uint32_t h, const float light[3],
                            const float* albedo, float* image, uint32_t* hit_row, float* hit_shade, vx_stream stream)
{
    if (!s || !cam || !light || !albedo || !image || (!hit_row) != (!hit_shade))
        return vx::fail(VX_ERR_INVALID_ARG, "vx_albedo_forward: bad arguments");
    if (w == 0 || h == 0) return VX_OK;
    if (w > 65536 || h > 65536) return vx::fail(VX_ERR_INVALID_ARG, "vx_albedo_forward: image too large");
    VX_DEVICE_GUARD(s->device);
    if (vx_status us = vx::scene_usable(s, "vx_albedo"); us != VX_OK) return us;
    hipStream_t st = static_cast<hipStream_t>(stream);
    vx_status status;
    vx_albedo_state* a = vx::albedo_state(s, st, status);
    if (!a) return status;
    uint32_t lds = 0;
    const vx::SceneView v = vx::alb_view(s, lds);
    const vx::AlbCam c = vx::make_alb_cam(cam, light);
    const uint32_t tiles = ((w + 15) / 16) * ((h + 15) / 16);
    vx::k_albedo_forward<<<dim3(tiles), dim3(vx::kAlbBlock), lds, st>>>(v, c, w, h, a->base, albedo, image, hit_row,
                                                                        hit_shade);
    if (hipGetLastError() != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_albedo_forward: launch failed");
    return VX_OK;
}

} // extern "C"

namespace vx {

// vx_albedo_backward's body. The deterministic reduction's workspace comes from
// the scene (grown with hipMalloc) unless noalloc, when it is carved from the
// device's scratch array and the call allocates nothing.
static vx_status albedo_backward(vx_scene* s, const vx_camera* cam, uint32_t w, uint32_t h, const float light[3],
                                 const float* grad_image, const uint32_t* hit_row, const float* hit_shade,
                                 float* grad_albedo, uint32_t mode, hipStream_t st, bool noalloc)
{
    const char* where = noalloc ? "vx_albedo_backward_noalloc" : "vx_albedo_backward";
    if (!s || !cam || !light || !grad_image || !grad_albedo || (!hit_row) != (!hit_shade) || mode > 1)
        return fail(VX_ERR_INVALID_ARG, std::string(where) + ": bad arguments");
    if (w == 0 || h == 0) return VX_OK;
    if (w > 65536 || h > 65536) return fail(VX_ERR_INVALID_ARG, std::string(where) + ": image too large");
    const size_t px = (size_t)w * h;
    if (noalloc && px > VX_NOALLOC_MAX_PIXELS)
        return fail(VX_ERR_INVALID_ARG, std::string(where) + ": " + std::to_string(w) + "x" + std::to_string(h) +
                                            " is " + std::to_string(px) + " pixels, more than the " +
                                            std::to_string(VX_NOALLOC_MAX_PIXELS) + " its workspace holds");
    VX_DEVICE_GUARD(s->device);
    if (vx_status us = vx::scene_usable(s, "vx_albedo"); us != VX_OK) return us;
    vx_albedo_state* a = s->albedo;
    if (!a && noalloc)
        return fail(VX_ERR_INVALID_ARG, std::string(where) + ": the scene has no voxel row table yet (it is built, "
                                                             "with an allocation, by vx_scene_voxel_count or the first "
                                                             "vx_albedo_forward)");
    vx_status status = VX_OK;
    if (!a && !(a = albedo_state(s, st, status))) return status;
    uint32_t lds = 0;
    const SceneView v = alb_view(s, lds);
    const AlbCam c = make_alb_cam(cam, light);
    const uint32_t tiles = ((w + 15) / 16) * ((h + 15) / 16);
    const bool recompute = hit_row == nullptr;
    const dim3 grid(tiles), block(kAlbBlock);
    if (mode == VX_GRAD_ATOMIC) {
        if (recompute)
            k_albedo_adjoint<true, true><<<grid, block, lds, st>>>(v, c, w, h, a->base, grad_image, nullptr, nullptr,
                                                                   grad_albedo, nullptr, nullptr, nullptr, 0u);
        else
            k_albedo_adjoint<false, true><<<grid, block, 0, st>>>(v, c, w, h, a->base, grad_image, hit_row, hit_shade,
                                                                  grad_albedo, nullptr, nullptr, nullptr, 0u);
        if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": launch failed");
        return VX_OK;
    }
    // deterministic: sort (row, contribution) by row, select the run heads, sum each run in a fixed order
    if (px > 0x7FFFFFFFu) return fail(VX_ERR_INVALID_ARG, std::string(where) + ": deterministic mode needs w*h < 2^31");
    if (a->voxels >= 0xFFFFFFFFull) return fail(VX_ERR_UNSUPPORTED, std::string(where) + ": too many voxels");
    // misses get key V (the voxel count) and sort last; the sort looks at the key bits V needs
    const uint32_t miss_key = (uint32_t)a->voxels;
    int end_bit = 1;
    while (end_bit < 32 && (miss_key >> end_bit) != 0u) ++end_bit;
    size_t sort_bytes = 0, red_bytes = 0;
    const hipcub::CountingInputIterator<uint32_t> index(0u);
    if (rocprim::radix_sort_pairs<PairSortConfig>(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                  (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t)px, 0u,
                                                  (unsigned)end_bit, st) != hipSuccess ||
        hipcub::DeviceSelect::If(nullptr, red_bytes, index, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)px,
                                 RunHead{nullptr}, st) != hipSuccess)
        return fail(VX_ERR_HIP, std::string(where) + ": workspace query failed");
    DetWorkspace ws{};
    if (noalloc) {
        if (vx_status g = scratch_workspace(s->device, px, std::max(sort_bytes, red_bytes), ws); g != VX_OK) return g;
    } else {
        if (vx_status g = grow(a, px, std::max(sort_bytes, red_bytes)); g != VX_OK) return g;
        ws = a->ws;
    }
    if (recompute)
        k_albedo_adjoint<true, false><<<grid, block, lds, st>>>(v, c, w, h, a->base, grad_image, nullptr, nullptr,
                                                                nullptr, ws.keys, ws.vals, ws.pix, miss_key);
    else
        k_albedo_adjoint<false, false><<<grid, block, 0, st>>>(v, c, w, h, a->base, grad_image, hit_row, hit_shade,
                                                               nullptr, ws.keys, ws.vals, ws.pix, miss_key);
    if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": launch failed");
    size_t tb = ws.cap_tmp;
    if (rocprim::radix_sort_pairs<PairSortConfig>(ws.tmp, tb, ws.keys, ws.keys_sorted, ws.pix, ws.pix_sorted,
                                                  (uint32_t)px, 0u, (unsigned)end_bit, st) != hipSuccess)
        return fail(VX_ERR_HIP, std::string(where) + ": sort failed");
    tb = ws.cap_tmp;
    if (hipcub::DeviceSelect::If(ws.tmp, tb, index, ws.heads, ws.n_runs, (int)px, RunHead{ws.keys_sorted}, st) !=
        hipSuccess)
        return fail(VX_ERR_HIP, std::string(where) + ": run selection failed");
    // resident groups walk all runs (their count stays on the device); the keys
    // buffer, consumed by the sort, holds the list of long runs
    const unsigned groups = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)s->cu_count * 8u,
                                                                               (px + kAlbBlock - 1) / kAlbBlock));
    k_clear_long_count<<<1, 1, 0, st>>>(ws.n_runs);
    k_reduce_short_runs<<<dim3(groups), block, 0, st>>>(ws.keys_sorted, ws.vals, ws.pix_sorted, ws.heads, ws.n_runs,
                                                        (uint32_t)px, miss_key, grad_albedo, ws.keys);
    k_reduce_long_runs<<<dim3(groups), block, 0, st>>>(ws.keys_sorted, ws.vals, ws.pix_sorted, ws.heads, ws.n_runs,
                                                       (uint32_t)px, ws.keys, grad_albedo);
    if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": run reduction failed");
    return VX_OK;
}

// At vx_ctx_create, with the device current: the device address of its scratch
// array (each device holds its own copy), once per device.
void scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
}

} // namespace vx

extern "C" {

vx_status vx_albedo_backward(vx_scene* s, const vx_camera* cam, uint32_t w, uint32_t h, const float light[3],
                             const float* grad_image, const uint32_t* hit_row, const float* hit_shade,
                             float* grad_albedo, uint32_t mode, vx_stream stream)
{
    return vx::albedo_backward(s, cam, w, h, light, grad_image, hit_row, hit_shade, grad_albedo, mode,
                               static_cast<hipStream_t>(stream), false);
}

vx_status vx_albedo_backward_noalloc(vx_scene* s, const vx_camera* cam, uint32_t w, uint32_t h, const float light[3],
                                     const float* grad_image, const uint32_t* hit_row, const float* hit_shade,
                                     float* grad_albedo, uint32_t mode, vx_stream stream)
{
    return vx::albedo_backward(s, cam, w, h, light, grad_image, hit_row, hit_shade, grad_albedo, mode,
                               static_cast<hipStream_t>(stream), true);
}

} // extern "C"
