// vx_cache.cpp -- VIXEN's on-disk scene cache (VSCA version 1): reader and
// writer, so scenes VIXEN cached load here byte for byte and scenes built here
// load in VIXEN.
//
// Layout (VoxelSceneCacher::SerializeToFile / DeserializeFromFile,
// libraries/CashSystem/src/VoxelSceneCacher.cpp:179-364; types from
// libraries/CashSystem/include/VoxelSceneCacher.h:38-44,145-173,190-227),
// little-endian, no padding:
//   u32 magic 0x56534341, u32 version 1, u32 entry count, then per entry:
//   u64 key = ((sceneType * 31 + resolution) * 31 + u32(density * 100)) * 31 + seed
//   u32 sceneType, u32 resolution, f32 density, u32 seed
//   six vectors, each u64 element count + elements:
//     esvoNodes (bytes), brickData (bytes), materials (32-byte GPUMaterial),
//     compressedColors (bytes), compressedNormals (bytes), brickGridLookup (u32)
//   OctreeConfig, sizeof(OctreeConfig) = 260 bytes: the 256-byte UBO, then
//     f32 worldGridSize at +256 (VoxelSceneCacher.h:100-133, written whole at
//     VoxelSceneCacher.cpp:254 and read whole at :336)
//   u32 nodeCount, u32 brickCount, u32 solidVoxelCount, u32 resolution, u32 sceneType
#include "vixen_build.h"
#include "vx_common.hpp"

#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

struct vx_cache {
    struct Entry {
        uint64_t key;
        uint32_t scene_type, resolution, seed;
        float density;
        std::vector<uint8_t> nodes, bricks, materials, colors, normals;
        std::vector<uint32_t> lookup;
        vx_octree_config cfg;
        float world_grid_size;
        uint32_t node_count, brick_count, solid_voxels, meta_resolution, meta_scene_type;
    };
    std::vector<Entry> entries;
};

namespace vx {
namespace {

constexpr uint32_t kMagic = 0x56534341u;
constexpr uint32_t kVersion = 1u;
constexpr uint64_t kMaterialBytes = 32;
// VoxelSceneCacher::BuildOctree's WORLD_GRID_SIZE (VoxelSceneCacher.cpp:605-607)
constexpr float kWorldGridSize = 10.0f;
static_assert(sizeof(vx_octree_config) == 256, "the OctreeConfig UBO portion is 256 bytes");

uint64_t cache_key(uint32_t scene_type, uint32_t resolution, float density, uint32_t seed)
{
    const uint32_t dq = static_cast<uint32_t>(density * 100.0f);
    uint64_t h = scene_type;
    h = h * 31 + resolution;
    h = h * 31 + dq;
    h = h * 31 + seed;
    return h;
}

struct Reader {
    const uint8_t* p;
    size_t left;
    template <class T>
    bool get(T& v)
    {
        if (left < sizeof(T)) return false;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        left -= sizeof(T);
        return true;
    }
    template <class T>
    bool vec(std::vector<T>& out, uint64_t elem_bytes = sizeof(T))
    {
        uint64_t n = 0;
        if (!get(n) || n > left / elem_bytes) return false;
        out.resize(n * elem_bytes / sizeof(T));
        if (n) std::memcpy(out.data(), p, n * elem_bytes);
        p += n * elem_bytes;
        left -= n * elem_bytes;
        return true;
    }
};

struct Writer {
    std::vector<uint8_t> buf;
    template <class T>
    void put(const T& v)
    {
        const auto* b = reinterpret_cast<const uint8_t*>(&v);
        buf.insert(buf.end(), b, b + sizeof(T));
    }
    void vec(const void* data, uint64_t count, uint64_t elem_bytes)
    {
        put(count);
        if (count) {
            const auto* b = static_cast<const uint8_t*>(data);
            buf.insert(buf.end(), b, b + count * elem_bytes);
        }
    }
};

vx_status read_file(const char* path, std::vector<uint8_t>& bytes)
{
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return fail(VX_ERR_INVALID_ARG, std::string("vx_cache_read: cannot open ") + path);
    std::vector<uint8_t> out;
    uint8_t chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0) out.insert(out.end(), chunk, chunk + n);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) return fail(VX_ERR_INVALID_ARG, std::string("vx_cache_read: read error on ") + path);
    bytes.swap(out);
    return VX_OK;
}

} // namespace
} // namespace vx

extern "C" {

static vx_status cache_read_impl(const char* path, vx_cache** out)
{
    if (!path || !out) return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_read: null argument");
    *out = nullptr;
    std::vector<uint8_t> bytes;
    if (vx_status st = vx::read_file(path, bytes); st != VX_OK) return st;
    vx::Reader r{bytes.data(), bytes.size()};
    uint32_t magic = 0, version = 0, count = 0;
    if (!r.get(magic) || !r.get(version)) return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: truncated header");
    if (magic != vx::kMagic) return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: not a VSCA scene cache (bad magic)");
    if (version != vx::kVersion)
        return vx::fail(VX_ERR_UNSUPPORTED, "vx_cache_read: cache version " + std::to_string(version) +
                                                " (stale; this reader knows version 1)");
    if (!r.get(count)) return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: truncated header");
    auto* c = new (std::nothrow) vx_cache{};
    if (!c) return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_cache_read: out of memory");
    for (uint32_t i = 0; i < count; ++i) {
        vx_cache::Entry e{};
        const std::string at = " (entry " + std::to_string(i) + ")";
        bool ok = r.get(e.key) && r.get(e.scene_type) && r.get(e.resolution) && r.get(e.density) && r.get(e.seed);
        if (ok && vx::cache_key(e.scene_type, e.resolution, e.density, e.seed) != e.key) {
            delete c;
            return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: key does not match the create info" + at);
        }
        ok = ok && r.vec(e.nodes) && r.vec(e.bricks) && r.vec(e.materials, vx::kMaterialBytes) && r.vec(e.colors) &&
             r.vec(e.normals) && r.vec(e.lookup) && r.get(e.cfg) && r.get(e.world_grid_size) &&
             r.get(e.node_count) && r.get(e.brick_count) && r.get(e.solid_voxels) && r.get(e.meta_resolution) &&
             r.get(e.meta_scene_type);
        if (!ok) {
            delete c;
            return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: truncated file" + at);
        }
        const uint64_t nb = e.brick_count;
        if (e.nodes.size() != uint64_t{e.node_count} * 8 || e.bricks.size() != nb * 2048 ||
            (!e.colors.empty() && e.colors.size() != nb * 256) || (!e.normals.empty() && e.normals.size() != nb * 512)) {
            delete c;
            return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: buffer sizes disagree with node/brick counts" + at);
        }
        c->entries.push_back(std::move(e));
    }
    *out = c;
    return VX_OK;
}

vx_status vx_cache_entry_count(const vx_cache* cache, uint32_t* count)
{
    if (!cache || !count) return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_entry_count: null argument");
    *count = static_cast<uint32_t>(cache->entries.size());
    return VX_OK;
}

vx_status vx_cache_entry(const vx_cache* cache, uint32_t index, vx_cache_entry_info* info, vx_scene_desc* desc)
{
    if (!cache || index >= cache->entries.size())
        return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_entry: no such entry");
    const vx_cache::Entry& e = cache->entries[index];
    if (info) {
        info->key = e.key;
        info->scene_type = e.scene_type;
        info->resolution = e.resolution;
        info->density = e.density;
        info->seed = e.seed;
        info->node_count = e.node_count;
        info->brick_count = e.brick_count;
        info->solid_voxel_count = e.solid_voxels;
        info->materials = e.materials.empty() ? nullptr : e.materials.data();
        info->material_count = e.materials.size() / vx::kMaterialBytes;
        info->brick_grid_lookup = e.lookup.empty() ? nullptr : e.lookup.data();
        info->brick_grid_lookup_count = e.lookup.size();
        info->world_grid_size = e.world_grid_size;
    }
    if (desc) {
        desc->esvo_nodes = reinterpret_cast<const uint32_t*>(e.nodes.data());
        desc->node_count = e.node_count;
        desc->brick_data = e.bricks.empty() ? nullptr : reinterpret_cast<const uint32_t*>(e.bricks.data());
        desc->brick_count = e.brick_count;
        desc->bc1_colors = e.colors.empty() ? nullptr : reinterpret_cast<const uint64_t*>(e.colors.data());
        desc->config = &e.cfg;
        desc->dxt_normals = e.normals.empty() ? nullptr : reinterpret_cast<const uint64_t*>(e.normals.data());
    }
    return VX_OK;
}

vx_status vx_cache_free(vx_cache* cache)
{
    delete cache;
    return VX_OK;
}

static vx_status cache_write_impl(const char* path, const vx_cache_entry_in* entries, uint32_t n)
{
    if (!path || (n && !entries)) return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_write: null argument");
    vx::Writer w;
    w.put(vx::kMagic);
    w.put(vx::kVersion);
    w.put(n);
    for (uint32_t i = 0; i < n; ++i) {
        const vx_cache_entry_in& e = entries[i];
        const vx_scene_desc* d = e.scene;
        if (!d || !d->esvo_nodes || !d->config || (d->brick_count && !d->brick_data) ||
            d->node_count > 0xFFFFFFFFull || d->brick_count > 0xFFFFFFFFull)
            return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_write: bad scene in entry " + std::to_string(i));
        const uint64_t nb = d->brick_count;
        w.put(vx::cache_key(e.scene_type, e.resolution, e.density, e.seed));
        w.put(e.scene_type);
        w.put(e.resolution);
        w.put(e.density);
        w.put(e.seed);
        w.vec(d->esvo_nodes, d->node_count * 8, 1);
        w.vec(d->brick_data, nb * 2048, 1);
        w.vec(e.materials, e.materials ? e.material_count : 0, vx::kMaterialBytes);
        w.vec(d->bc1_colors, d->bc1_colors ? nb * 256 : 0, 1);
        w.vec(d->dxt_normals, d->dxt_normals ? nb * 512 : 0, 1);
        w.vec(e.brick_grid_lookup, e.brick_grid_lookup ? e.brick_grid_lookup_count : 0, 4);
        w.put(*d->config);
        const float wgs = e.world_grid_size;
        w.put(wgs > 0.0f && wgs < __builtin_inff() ? wgs : vx::kWorldGridSize); // sizeof(OctreeConfig) = 260
        w.put(static_cast<uint32_t>(d->node_count));
        w.put(static_cast<uint32_t>(nb));
        w.put(e.solid_voxel_count);
        w.put(e.resolution);
        w.put(e.scene_type);
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return vx::fail(VX_ERR_INVALID_ARG, std::string("vx_cache_write: cannot create ") + path);
    const bool ok = std::fwrite(w.buf.data(), 1, w.buf.size(), f) == w.buf.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return vx::fail(VX_ERR_INVALID_ARG, std::string("vx_cache_write: write failed on ") + path);
    return VX_OK;
}

// Nothing may throw across the ABI: allocation failures become statuses.
vx_status vx_cache_read(const char* path, vx_cache** out)
{
    try {
        return cache_read_impl(path, out);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_cache_read: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_SCENE, "vx_cache_read: unexpected exception");
    }
}

vx_status vx_cache_write(const char* path, const vx_cache_entry_in* entries, uint32_t n)
{
    try {
        return cache_write_impl(path, entries, n);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_cache_write: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_ARG, "vx_cache_write: unexpected exception");
    }
}

} // extern "C"
