// vx_svo.cpp -- the C++ ISVOStructure-shaped adapter of vixen_svo.hpp: host
// copies of the scene buffers with LaineKarrasOctree's structure queries, and
// the castRay family over vx_trace / vx_trace_lod.
#include "vixen_svo.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <system_error>
#include <thread>

namespace vx::svo {

namespace {

[[noreturn]] void raise(const std::string& what) { throw std::runtime_error(what + ": " + vx_last_error()); }

void check(vx_status st, const char* what)
{
    if (st != VX_OK) raise(what);
}

void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Makes `device` current for one call and gives the caller's device back.
class OnDevice {
public:
    explicit OnDevice(int device)
    {
        check_hip(hipGetDevice(&prev_), "hipGetDevice");
        if (prev_ != device) check_hip(hipSetDevice(device), "hipSetDevice");
    }
    ~OnDevice() { (void)hipSetDevice(prev_); }
    OnDevice(const OnDevice&) = delete;
    OnDevice& operator=(const OnDevice&) = delete;

private:
    int prev_ = 0;
};

vx_lod to_lod(const LODParameters& p) { return vx_lod{p.rayOrigSize, p.rayDirSize}; }
LODParameters from_lod(const vx_lod& l) { return LODParameters{l.ray_orig_size, l.ray_dir_size}; }

// Rays per chunk of the host-memory casts (2 MiB of rays and 2 MiB of hits).
constexpr size_t kStage = size_t{1} << 16;

} // namespace

LODParameters LODParameters::fromCamera(float fovY, int screenHeight)
{
    return from_lod(vx_lod_from_camera(fovY, screenHeight));
}

LODParameters LODParameters::fromCameraWithNearPlane(float fovY, int screenHeight, float nearPlane)
{
    return from_lod(vx_lod_from_camera_near(fovY, screenHeight, nearPlane));
}

LODParameters LODParameters::withBias(float bias) const { return from_lod(vx_lod_with_bias(to_lod(*this), bias)); }

// ---------------------------------------------------------------------------
// HostOctree
// ---------------------------------------------------------------------------

HostOctree::HostOctree(const vx_scene_desc& d)
{
    check(vx_scene_validate(&d), "HostOctree");
    nodes_.assign(d.esvo_nodes, d.esvo_nodes + 2 * d.node_count);
    bricks_.assign(d.brick_data, d.brick_data + 512 * d.brick_count);
    if (d.bc1_colors) bc1_.assign(d.bc1_colors, d.bc1_colors + 32 * d.brick_count);
    if (d.dxt_normals) normals_.assign(d.dxt_normals, d.dxt_normals + 64 * d.brick_count);
    cfg_ = *d.config;
    voxel_count_ = static_cast<size_t>(std::count_if(bricks_.begin(), bricks_.end(), [](uint32_t w) { return w != 0; }));
}

vx_scene_desc HostOctree::desc() const
{
    vx_scene_desc d{};
    d.esvo_nodes = nodes_.data();
    d.node_count = nodes_.size() / 2;
    d.brick_data = bricks_.data();
    d.brick_count = bricks_.size() / 512;
    d.bc1_colors = bc1_.empty() ? nullptr : bc1_.data();
    d.dxt_normals = normals_.empty() ? nullptr : normals_.data();
    d.config = &cfg_;
    return d;
}

Vec3 HostOctree::getWorldMin() const { return Vec3{cfg_.grid_min[0], cfg_.grid_min[1], cfg_.grid_min[2]}; }
Vec3 HostOctree::getWorldMax() const { return Vec3{cfg_.grid_max[0], cfg_.grid_max[1], cfg_.grid_max[2]}; }

VoxelBounds HostOctree::getVoxelBounds(const Vec3&, int) const { return VoxelBounds{getWorldMin(), getWorldMax()}; }

float HostOctree::getVoxelSize(int scale) const
{
    if (scale >= getMaxLevels()) return 0.0f;
    const float ext = cfg_.grid_max[0] - cfg_.grid_min[0];
    return ext / std::pow(2.0f, static_cast<float>(scale));
}

size_t HostOctree::getMemoryUsage() const
{
    const size_t nb = bricks_.size() / 512;
    return nodes_.size() * 4 + bricks_.size() * 4 + bc1_.size() * 8 + normals_.size() * 8 + nb * 68;
}

std::string HostOctree::getStats() const
{
    const size_t voxels = getVoxelCount(), mem = getMemoryUsage();
    std::ostringstream oss;
    oss << "Laine-Karras SVO Statistics:\n";
    oss << "  Total voxels: " << voxels << "\n";
    oss << "  Max levels: " << getMaxLevels() << "\n";
    oss << "  Memory usage: " << (mem / 1024.0 / 1024.0) << " MB\n";
    oss << "  Avg bytes/voxel: " << (voxels > 0 ? mem / voxels : 0) << "\n";
    return oss.str();
}

GPUBuffers HostOctree::getGPUBuffers() const
{
    auto bytes = [](const auto& v) {
        const auto* p = reinterpret_cast<const uint8_t*>(v.data());
        return std::vector<uint8_t>(p, p + v.size() * sizeof(v[0]));
    };
    GPUBuffers b;
    b.hierarchyBuffer = bytes(nodes_);
    b.attributeBuffer = bytes(bricks_);
    b.compressedColorBuffer = bytes(bc1_);
    b.compressedNormalBuffer = bytes(normals_);
    return b;
}

// Child `octant` of descriptor `node`: the node's child block (the 15-bit
// childPointer, or word1 bits 0-23 under farBit) plus the non-leaf children
// before it, as LaineKarrasOctree.cpp:152-165 counts them.
uint32_t HostOctree::child_index(uint32_t node, int octant) const
{
    const uint32_t w0 = nodes_[2 * size_t{node}], w1 = nodes_[2 * size_t{node} + 1];
    const uint32_t valid = (w0 >> 16) & 0xFFu, leaf = (w0 >> 24) & 0xFFu;
    const uint32_t block = (w0 & 0x8000u) ? (w1 & 0xFFFFFFu) : (w0 & 0x7FFFu);
    uint32_t before = 0;
    for (int i = 0; i < octant; ++i)
        if ((valid >> i & 1u) && !(leaf >> i & 1u)) ++before;
    return block + before;
}

// The descent both queries share (LaineKarrasOctree.cpp:108-170, 259-314):
// -1 outside the world box, 0 at an empty child, 1 at a leaf (brick) child,
// 2 after `scale` levels, with `node` the descriptor reached.
int HostOctree::descend(const Vec3& p, int scale, uint32_t& node) const
{
    node = 0;
    if (nodes_.empty()) return -1;
    const Vec3 lo = getWorldMin(), hi = getWorldMax();
    if (p.x < lo.x || p.y < lo.y || p.z < lo.z || p.x >= hi.x || p.y >= hi.y || p.z >= hi.z) return -1;
    const float nx = (p.x - lo.x) / (hi.x - lo.x), ny = (p.y - lo.y) / (hi.y - lo.y), nz = (p.z - lo.z) / (hi.z - lo.z);
    float px = 0.0f, py = 0.0f, pz = 0.0f, size = 1.0f;
    for (int level = 0; level < scale; ++level) {
        size *= 0.5f;
        int idx = 0;
        float cx = px, cy = py, cz = pz;
        if (nx >= px + size) {
            idx |= 1;
            cx += size;
        }
        if (ny >= py + size) {
            idx |= 2;
            cy += size;
        }
        if (nz >= pz + size) {
            idx |= 4;
            cz += size;
        }
        const uint32_t w0 = nodes_[2 * size_t{node}];
        if (!((w0 >> 16) >> idx & 1u)) return 0;
        if ((w0 >> 24) >> idx & 1u) return 1;
        node = child_index(node, idx);
        px = cx;
        py = cy;
        pz = cz;
    }
    return 2;
}

bool HostOctree::voxelExists(const Vec3& position, int scale) const
{
    uint32_t node;
    const int r = descend(position, scale, node);
    return r == 1 || r == 2;
}

uint8_t HostOctree::getChildMask(const Vec3& position, int scale) const
{
    uint32_t node;
    if (descend(position, scale, node) != 2) return 0;
    return static_cast<uint8_t>((nodes_[2 * size_t{node}] >> 16) & 0xFFu);
}

RayHit HostOctree::toRayHit(const vx_ray& r, const vx_hit& h) const
{
    RayHit o;
    o.brick = h.brick;
    o.voxel = h.voxel;
    o.leafNode = h.leaf_node;
    o.iterations = VX_HIT_ITERS(h.flags);
    if (!(h.flags & VX_HIT_FLAG_HIT)) return o; // the reference's value-initialised miss
    // glm::normalize: d * (1 / sqrt(dot(d, d))), dot summed x, y, then z
    const float len2 = (r.d[0] * r.d[0] + r.d[1] * r.d[1]) + r.d[2] * r.d[2];
    const float inv = 1.0f / std::sqrt(len2);
    const float dx = r.d[0] * inv, dy = r.d[1] * inv, dz = r.d[2] * inv;
    const float t = h.t;
    o.hit = true;
    o.tMin = t;
    o.hitPoint = Vec3{r.o[0] + dx * t, r.o[1] + dy * t, r.o[2] + dz * t};
    o.color = Vec3{h.rgb[0], h.rgb[1], h.rgb[2]};
    const float ext = cfg_.grid_max[0] - cfg_.grid_min[0];
    if (h.flags & VX_HIT_FLAG_LOD) {
        o.lod = true;
        o.scale = static_cast<int>(VX_HIT_LOD_LEVEL(h.flags));
        o.normal = Vec3{0.0f, 1.0f, 0.0f};
        // esvoScaleToWorldSize of the child's ESVO scale, user level + 23 - maxLevels
        o.tMax = t + std::ldexp(1.0f, o.scale - getMaxLevels()) * ext;
    } else {
        o.scale = getMaxLevels() - 1;
        const uint32_t axis = VX_HIT_AXIS(h.flags);
        const float sg = (h.flags & 8u) ? 1.0f : -1.0f;
        o.normal = Vec3{axis == 0 ? sg : 0.0f, axis == 1 ? sg : 0.0f, axis == 2 ? sg : 0.0f};
        o.tMax = t + ext / static_cast<float>(1u << getMaxLevels()); // one voxel's edge
    }
    return o;
}

// ---------------------------------------------------------------------------
// LaineKarrasOctreeHip
// ---------------------------------------------------------------------------

LaineKarrasOctreeHip::LaineKarrasOctreeHip(int device, const vx_scene_desc& d) : HostOctree(d), device_(device)
{
    check(vx_ctx_create(device, &ctx_), "LaineKarrasOctreeHip: vx_ctx_create");
    const vx_scene_desc own = desc();
    if (vx_scene_create(ctx_, &own, &scene_) != VX_OK) {
        const std::string why = vx_last_error();
        (void)vx_ctx_destroy(ctx_);
        throw std::runtime_error("LaineKarrasOctreeHip: vx_scene_create: " + why);
    }
}

LaineKarrasOctreeHip::~LaineKarrasOctreeHip()
{
    // no exceptions here: the device is switched by hand, and every call's status is dropped
    int prev = -1;
    const bool switched = hipGetDevice(&prev) == hipSuccess && prev != device_ && hipSetDevice(device_) == hipSuccess;
    for (Slot& sl : slots_) {
        if (sl.stream) {
            (void)hipStreamSynchronize(static_cast<hipStream_t>(sl.stream));
            (void)hipStreamDestroy(static_cast<hipStream_t>(sl.stream));
        }
        (void)hipFree(sl.d_rays);
        (void)hipFree(sl.d_hits);
        (void)hipHostFree(sl.h_rays);
        (void)hipHostFree(sl.h_hits);
    }
    if (switched) (void)hipSetDevice(prev);
    (void)vx_scene_destroy(scene_);
    (void)vx_ctx_destroy(ctx_);
}

vx_status LaineKarrasOctreeHip::castRaysDevice(const vx_ray* rays, uint64_t n, vx_hit* hits, vx_stream stream,
                                               const LODParameters* lod) const
{
    if (lod) {
        const vx_lod l = to_lod(*lod);
        return vx_trace_lod(scene_, rays, n, &l, hits, stream);
    }
    return vx_trace(scene_, rays, n, hits, stream);
}

void LaineKarrasOctreeHip::trace_host(const vx_ray* rays, size_t n, const LODParameters* lod, vx_hit* out) const
{
    std::lock_guard<std::mutex> lock(mu_);
    OnDevice on(device_);
    for (Slot& sl : slots_) {
        if (!sl.stream) {
            hipStream_t st = nullptr;
            check_hip(hipStreamCreate(&st), "hipStreamCreate");
            sl.stream = st;
        }
        // a cast that threw may have left copies in flight on the slot
        check_hip(hipStreamSynchronize(static_cast<hipStream_t>(sl.stream)), "hipStreamSynchronize");
        if (!sl.d_rays) check_hip(hipMalloc(reinterpret_cast<void**>(&sl.d_rays), kStage * sizeof(vx_ray)), "hipMalloc");
        if (!sl.d_hits) check_hip(hipMalloc(reinterpret_cast<void**>(&sl.d_hits), kStage * sizeof(vx_hit)), "hipMalloc");
        if (!sl.h_rays)
            check_hip(hipHostMalloc(reinterpret_cast<void**>(&sl.h_rays), kStage * sizeof(vx_ray), hipHostMallocDefault),
                      "hipHostMalloc");
        if (!sl.h_hits)
            check_hip(hipHostMalloc(reinterpret_cast<void**>(&sl.h_hits), kStage * sizeof(vx_hit), hipHostMallocDefault),
                      "hipHostMalloc");
    }
    // chunk c uses slot c % 2: its rays are copied into the slot's page-locked buffer, sent, traced and sent
    // back on the slot's stream while the host fills the other slot; a slot's records are copied out before
    // the slot takes its next chunk
    const size_t chunks = (n + kStage - 1) / kStage;
    auto retire = [&](size_t c) {
        const Slot& sl = slots_[c & 1u];
        const size_t at = c * kStage, m = std::min(kStage, n - at);
        check_hip(hipStreamSynchronize(static_cast<hipStream_t>(sl.stream)), "hipStreamSynchronize");
        std::memcpy(out + at, sl.h_hits, m * sizeof(vx_hit));
    };
    for (size_t c = 0; c < chunks; ++c) {
        if (c >= 2) retire(c - 2);
        const Slot& sl = slots_[c & 1u];
        const auto st = static_cast<hipStream_t>(sl.stream);
        const size_t at = c * kStage, m = std::min(kStage, n - at);
        std::memcpy(sl.h_rays, rays + at, m * sizeof(vx_ray));
        check_hip(hipMemcpyAsync(sl.d_rays, sl.h_rays, m * sizeof(vx_ray), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        check(castRaysDevice(sl.d_rays, m, sl.d_hits, st, lod), "castRays");
        check_hip(hipMemcpyAsync(sl.h_hits, sl.d_hits, m * sizeof(vx_hit), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    }
    for (size_t c = chunks >= 2 ? chunks - 2 : 0; c < chunks; ++c) retire(c);
}

std::vector<RayHit> LaineKarrasOctreeHip::castRays(const vx_ray* rays, size_t n, const LODParameters* lod) const
{
    std::vector<vx_hit> h(n);
    if (n) trace_host(rays, n, lod, h.data());
    std::vector<RayHit> out(n);
    // the RayHit conversion over up to 16 host threads for large batches (a range whose thread cannot be
    // started runs inline)
    const auto convert = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) out[i] = toRayHit(rays[i], h[i]);
    };
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = n < (size_t{1} << 16) ? 1 : std::min<size_t>(std::min<size_t>(hw, 16), n >> 14);
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; ++t) {
        const size_t b = n * t / nt, e = n * (t + 1) / nt;
        try {
            pool.emplace_back(convert, b, e);
        } catch (const std::system_error&) {
            convert(b, e);
        }
    }
    convert(0, n / nt);
    for (std::thread& th : pool) th.join();
    return out;
}

namespace {
vx_ray one_ray(const Vec3& o, const Vec3& d, float tMin, float tMax)
{
    return vx_ray{{o.x, o.y, o.z}, tMin, {d.x, d.y, d.z}, tMax};
}
} // namespace

RayHit LaineKarrasOctreeHip::castRay(const Vec3& origin, const Vec3& direction, float tMin, float tMax) const
{
    const vx_ray r = one_ray(origin, direction, tMin, tMax);
    return castRays(&r, 1)[0];
}

RayHit LaineKarrasOctreeHip::castRayLOD(const Vec3& origin, const Vec3& direction, float lodBias, float tMin,
                                        float tMax) const
{
    LODParameters p;
    p.rayDirSize = lodBias;
    p.rayOrigSize = 0.0f;
    return castRayWithLOD(origin, direction, p, tMin, tMax);
}

RayHit LaineKarrasOctreeHip::castRayScreenSpaceLOD(const Vec3& origin, const Vec3& direction, float fovY,
                                                   int screenHeight, float tMin, float tMax) const
{
    return castRayWithLOD(origin, direction, LODParameters::fromCamera(fovY, screenHeight), tMin, tMax);
}

RayHit LaineKarrasOctreeHip::castRayWithLOD(const Vec3& origin, const Vec3& direction, const LODParameters& lodParams,
                                            float tMin, float tMax) const
{
    const vx_ray r = one_ray(origin, direction, tMin, tMax);
    return castRays(&r, 1, &lodParams)[0];
}

} // namespace vx::svo
