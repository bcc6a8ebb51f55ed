/*
 * vixen_svo.hpp -- C++ adapter with the ray-cast surface of VIXEN's
 * ISVOStructure / LaineKarrasOctree (libraries/SVO/include/ISVOStructure.h:36-181,
 * libraries/SVO/include/LaineKarrasOctree.h:64-108), over the C ABI of
 * vixen_hip.h. A C++ caller that casts rays through an ISVOStructure switches
 * to LaineKarrasOctreeHip and keeps its call sites: the method names, argument
 * meaning and RayHit fields are the reference's, the traversal runs on the GPU.
 *
 * Frame. The adapter's world is the box the scene's config names (grid_min ..
 * grid_max), traced as it is: VIXEN's CPU octree traces in voxel units [0,N]^3
 * (VoxelSceneCacher.cpp:526-537), where local and world coincide.
 *
 * RayHit as the reference builds it (SVOBrickDDA.cpp:505-511 for a voxel hit,
 * SVOTraversal.cpp:637-647 for an LOD stop, a value-initialised record for a
 * miss, SVOTraversal.cpp:524-526):
 *   voxel hit: hit, tMin = entry t of the hit voxel, tMax = tMin + the voxel's
 *              edge (1 in voxel units, the reference's brickVoxelSize), hitPoint
 *              = origin + normalize(direction) * tMin, normal = the entry face's
 *              (-sign(direction) on the entry axis), scale = getMaxLevels() - 1;
 *   LOD stop:  hit, tMin = world distance of the coarse child's entry, tMax =
 *              tMin + its edge (esvoScaleToWorldSize), hitPoint as above, normal
 *              (0,1,0), scale = its user level;
 *   miss:      every field zero.
 * entity is VIXEN's gaia ECS handle; with no ECS here it is 0 (the reference's
 * gaia::ecs::Entity() for hits without one), and brick / voxel / leafNode name
 * the hit voxel instead (vx_hit's fields).
 *
 * Errors: the constructors throw std::runtime_error (with vx_last_error's text);
 * the ray casts themselves have no error path in the reference, so a failed GPU
 * call throws there too.
 */
#ifndef VIXEN_SVO_HPP
#define VIXEN_SVO_HPP

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "vixen_hip.h"

namespace vx::svo {

/* glm::vec3's layout: three floats. */
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

/* LODParameters (libraries/SVO/include/SVOLOD.h:50-195): the ray cone. */
struct LODParameters {
    float rayOrigSize = 0.0f;
    float rayDirSize = 0.0f;
    static LODParameters fromCamera(float fovY, int screenHeight);
    static LODParameters fromCameraWithNearPlane(float fovY, int screenHeight, float nearPlane);
    LODParameters withBias(float bias) const;
    bool isEnabled() const { return rayDirSize > 0.0f; }
};

/* ISVOStructure::RayHit (ISVOStructure.h:101-112) plus the GPU record behind it. */
struct RayHit {
    uint64_t entity = 0;
    Vec3 hitPoint;
    Vec3 normal;
    float tMin = 0.0f;
    float tMax = 0.0f;
    int scale = 0;
    bool hit = false;
    // vx_hit's fields (~0u where the record has none)
    uint32_t brick = ~0u;
    uint32_t voxel = ~0u;
    uint32_t leafNode = ~0u;
    Vec3 color;
    uint32_t iterations = 0;
    bool lod = false;
};

/* ISVOStructure::VoxelBounds (ISVOStructure.h:67-73). */
struct VoxelBounds {
    Vec3 min;
    Vec3 max;
};

/* ISVOStructure::GPUBuffers (ISVOStructure.h:252-269). */
struct GPUBuffers {
    std::vector<uint8_t> hierarchyBuffer;
    std::vector<uint8_t> attributeBuffer;
    std::vector<uint8_t> auxBuffer;
    std::vector<uint8_t> compressedColorBuffer;
    std::vector<uint8_t> compressedNormalBuffer;
};

/* The host side of the structure: a copy of the scene buffers and the
 * queries that read them (no GPU involved). */
class HostOctree {
public:
    explicit HostOctree(const vx_scene_desc& desc); // copies the buffers; throws on an invalid scene

    /* LaineKarrasOctree::voxelExists (LaineKarrasOctree.cpp:108-170): true when
     * the node of `scale` levels below the root that holds `position` is
     * populated, or a brick is reached on the way down. */
    bool voxelExists(const Vec3& position, int scale) const;
    /* LaineKarrasOctree::getChildMask (LaineKarrasOctree.cpp:259-314): that
     * node's validMask, 0 when it is absent or a brick is reached first. */
    uint8_t getChildMask(const Vec3& position, int scale) const;
    /* The reference returns the world box for every voxel (LaineKarrasOctree.cpp:316-321). */
    VoxelBounds getVoxelBounds(const Vec3& position, int scale) const;

    Vec3 getWorldMin() const;
    Vec3 getWorldMax() const;
    int getMaxLevels() const { return cfg_.user_max_levels; }
    /* LaineKarrasOctree::getVoxelSize (LaineKarrasOctree.cpp:327-334). */
    float getVoxelSize(int scale) const;
    /* Solid voxels (brickData words != 0), the rebuild's totalVoxels (SVORebuild.cpp:281-318). */
    size_t getVoxelCount() const { return voxel_count_; }
    /* Bytes of the scene buffers held on the GPU (descriptors, brickData, BC1 and
     * normal pools, and the per-brick occupancy and box tables vx_scene_create adds). */
    size_t getMemoryUsage() const;
    /* LaineKarrasOctree::getStats's text (LaineKarrasOctree.cpp:336-348). */
    std::string getStats() const;
    /* The scene buffers as bytes: hierarchy = the child descriptors, attribute =
     * brickData, and the BC1 colour and DXT normal pools (empty when absent). The
     * reference packs only the two compressed pools (LaineKarrasOctree.cpp:369-398;
     * its hierarchy, attribute and aux buffers are left empty, a TODO at :376). */
    GPUBuffers getGPUBuffers() const;

    const vx_octree_config& config() const { return cfg_; }
    /* vx_scene_desc pointing into this copy. */
    vx_scene_desc desc() const;

    /* A GPU record as the reference's RayHit for ray r (see the file header). */
    RayHit toRayHit(const vx_ray& r, const vx_hit& h) const;

private:
    uint32_t child_index(uint32_t node, int octant) const; // index of a child descriptor
    int descend(const Vec3& position, int scale, uint32_t& node) const;

    std::vector<uint32_t> nodes_;
    std::vector<uint32_t> bricks_;
    std::vector<uint64_t> bc1_;
    std::vector<uint64_t> normals_;
    vx_octree_config cfg_{};
    size_t voxel_count_ = 0;
};

/* LaineKarrasOctree's castRay surface on the GPU. */
class LaineKarrasOctreeHip : public HostOctree {
public:
    /* Uploads the scene (vx_scene_create) to GPU `device`, with a context of its own. */
    LaineKarrasOctreeHip(int device, const vx_scene_desc& desc);
    ~LaineKarrasOctreeHip();
    LaineKarrasOctreeHip(const LaineKarrasOctreeHip&) = delete;
    LaineKarrasOctreeHip& operator=(const LaineKarrasOctreeHip&) = delete;

    /* ISVOStructure.h:122-181 / SVOTraversal.cpp:189-232. */
    RayHit castRay(const Vec3& origin, const Vec3& direction, float tMin = 0.0f, float tMax = 1e30f) const;
    /* The legacy interface: lodBias is the cone's growth per unit distance (SVOTraversal.cpp:198-210). */
    RayHit castRayLOD(const Vec3& origin, const Vec3& direction, float lodBias, float tMin = 0.0f,
                      float tMax = 1e30f) const;
    RayHit castRayScreenSpaceLOD(const Vec3& origin, const Vec3& direction, float fovY, int screenHeight,
                                 float tMin = 0.0f, float tMax = 1e30f) const;
    RayHit castRayWithLOD(const Vec3& origin, const Vec3& direction, const LODParameters& lodParams,
                          float tMin = 0.0f, float tMax = 1e30f) const;

    /* Many rays at once from host memory (castRay, or castRayWithLOD when lod is
     * given), in 64 Ki-ray chunks through the adapter's two page-locked staging
     * slots (one chunk's transfers and trace overlap the host copies of the
     * next); large batches convert their records on up to 16 host threads. */
    std::vector<RayHit> castRays(const vx_ray* rays, size_t n, const LODParameters* lod = nullptr) const;
    /* The same on device buffers (vx_trace / vx_trace_lod on the caller's
     * stream; asynchronous, no copies). */
    vx_status castRaysDevice(const vx_ray* rays, uint64_t n, vx_hit* hits, vx_stream stream,
                             const LODParameters* lod = nullptr) const;

    vx_scene* scene() const { return scene_; }

private:
    void trace_host(const vx_ray* rays, size_t n, const LODParameters* lod, vx_hit* out) const;

    // Staging of the host-memory casts: two slots, each with device ray and hit
    // buffers, page-locked host copies of both and a stream, so that one chunk's
    // transfers and trace overlap the host-side copies of the next.
    struct Slot {
        vx_ray* d_rays = nullptr;
        vx_hit* d_hits = nullptr;
        vx_ray* h_rays = nullptr;
        vx_hit* h_hits = nullptr;
        void* stream = nullptr;
    };

    vx_ctx* ctx_ = nullptr;
    vx_scene* scene_ = nullptr;
    int device_ = 0;
    mutable std::mutex mu_;       // the staging slots serve one cast at a time
    mutable Slot slots_[2];
};

} // namespace vx::svo

#endif
