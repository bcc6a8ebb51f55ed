// svo_consumer.cpp -- a C++ host of the ISVOStructure-shaped adapter
// (include/vixen_svo.hpp), with no Python in its process. Built by the package
// build as tests/svo_consumer; the tests read what it writes.
//
//   svo_consumer queries OUT      host only: the structure queries on the
//                                 seeded 64^3 cave (seed 3, 30% occupancy)
//   svo_consumer convert IN OUT   host only: toRayHit of the rays and records
//                                 in IN (u64 n, n vx_ray, n vx_hit)
//   svo_consumer trace OUT        GPU 0: castRay, castRayLOD,
//                                 castRayScreenSpaceLOD, castRayWithLOD and
//                                 castRays on the seeded 128^3 cave with BC1
//                                 (seed 5, 50% occupancy)
//   svo_consumer rate N           GPU 0: castRays of N seeded rays from host
//                                 memory on the 256^3 cave (seed 2, 50%
//                                 occupancy), median of 5; prints one JSON line
//
// Each ray group is written as u64 n, n vx_ray, n records of 80 bytes:
// entity u64, hitPoint 3f, normal 3f, tMin f, tMax f, scale i32, hit u32,
// brick u32, voxel u32, leafNode u32, color 3f, iterations u32, lod u32.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "vixen_build.h"
#include "vixen_svo.hpp"

using vx::svo::HostOctree;
using vx::svo::LaineKarrasOctreeHip;
using vx::svo::LODParameters;
using vx::svo::RayHit;
using vx::svo::Vec3;

namespace {

struct Rec {
    uint64_t entity;
    float hp[3], n[3];
    float tmin, tmax;
    int32_t scale;
    uint32_t hit, brick, voxel, leaf;
    float color[3];
    uint32_t iters, lod;
};
static_assert(sizeof(Rec) == 80, "record layout");

Rec rec(const RayHit& h)
{
    return Rec{h.entity, {h.hitPoint.x, h.hitPoint.y, h.hitPoint.z}, {h.normal.x, h.normal.y, h.normal.z}, h.tMin,
               h.tMax, h.scale, h.hit ? 1u : 0u, h.brick, h.voxel, h.leafNode, {h.color.x, h.color.y, h.color.z},
               h.iterations, h.lod ? 1u : 0u};
}

struct Rng { // splitmix64
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform(float lo, float hi) { return lo + (hi - lo) * (float)(next() >> 40) * 0x1p-24f; }
};

struct Built {
    vx_built_scene* b = nullptr;
    ~Built() { vx_built_scene_free(b); }
    vx_scene_desc desc() const
    {
        vx_scene_desc d{};
        d.esvo_nodes = b->esvo_nodes;
        d.node_count = b->node_count;
        d.brick_data = b->brick_data;
        d.brick_count = b->brick_count;
        d.bc1_colors = b->bc1_colors;
        d.config = &b->config;
        return d;
    }
};

void build_cave(uint32_t N, uint64_t seed, double occ, bool bc1, Built& out)
{
    std::vector<uint32_t> grid(size_t{N} * N * N);
    vx_camera cam;
    if (vx_gen_cave(N, seed, occ, grid.data(), &cam) != VX_OK) throw std::runtime_error(vx_last_error());
    if (vx_build_from_dense(grid.data(), N, bc1 ? 1 : 0, 0, &out.b) != VX_OK) throw std::runtime_error(vx_last_error());
}

struct Writer {
    FILE* f;
    explicit Writer(const char* path) : f(std::fopen(path, "wb"))
    {
        if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    }
    ~Writer() { std::fclose(f); }
    void raw(const void* p, size_t n)
    {
        if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("short write");
    }
    template <class T> void put(const T& v) { raw(&v, sizeof(T)); }
    void group(const std::vector<vx_ray>& rays, const std::vector<RayHit>& hits)
    {
        put<uint64_t>(rays.size());
        raw(rays.data(), rays.size() * sizeof(vx_ray));
        for (const RayHit& h : hits) put(rec(h));
    }
};

int queries(const char* out)
{
    Built b;
    build_cave(64, 3, 0.3, false, b);
    const HostOctree oct(b.desc());
    Writer w(out);
    // scalar queries
    w.put<int32_t>(oct.getMaxLevels());
    w.put<uint64_t>(oct.getVoxelCount());
    w.put<uint64_t>(oct.getMemoryUsage());
    const Vec3 lo = oct.getWorldMin(), hi = oct.getWorldMax();
    w.put(lo);
    w.put(hi);
    for (int s = 0; s <= 12; ++s) w.put(oct.getVoxelSize(s));
    const vx::svo::VoxelBounds vb = oct.getVoxelBounds(Vec3{1, 2, 3}, 4);
    w.put(vb);
    const vx::svo::GPUBuffers g = oct.getGPUBuffers();
    w.put<uint64_t>(g.hierarchyBuffer.size());
    w.put<uint64_t>(g.attributeBuffer.size());
    w.put<uint64_t>(g.compressedColorBuffer.size());
    // positions: seeded points in and around the box, and points on cell corners
    Rng rng{17};
    std::vector<Vec3> pts;
    for (int i = 0; i < 20000; ++i) pts.push_back(Vec3{rng.uniform(-2, 66), rng.uniform(-2, 66), rng.uniform(-2, 66)});
    for (int i = 0; i < 4000; ++i)
        pts.push_back(Vec3{(float)(rng.next() % 65), (float)(rng.next() % 65), (float)(rng.next() % 65)});
    w.put<uint64_t>(pts.size());
    for (const Vec3& p : pts) {
        w.put(p);
        for (int s = 0; s <= 8; ++s) {
            w.put<uint8_t>(oct.voxelExists(p, s) ? 1 : 0);
            w.put<uint8_t>(oct.getChildMask(p, s));
        }
    }
    const std::string st = oct.getStats();
    w.put<uint64_t>(st.size());
    w.raw(st.data(), st.size());
    return 0;
}

int convert(const char* in, const char* out)
{
    Built b;
    build_cave(128, 5, 0.5, true, b);
    const HostOctree oct(b.desc());
    FILE* f = std::fopen(in, "rb");
    if (!f) throw std::runtime_error(std::string("cannot read ") + in);
    uint64_t n = 0;
    bool ok = std::fread(&n, 8, 1, f) == 1;
    std::vector<vx_ray> rays(n);
    std::vector<vx_hit> hits(n);
    ok = ok && std::fread(rays.data(), sizeof(vx_ray), n, f) == n && std::fread(hits.data(), sizeof(vx_hit), n, f) == n;
    std::fclose(f);
    if (!ok) throw std::runtime_error("short read");
    std::vector<RayHit> got(n);
    for (uint64_t i = 0; i < n; ++i) got[i] = oct.toRayHit(rays[i], hits[i]);
    Writer w(out);
    w.group(rays, got);
    return 0;
}

std::vector<vx_ray> seeded_rays(Rng& rng, size_t n, float N)
{
    std::vector<vx_ray> r(n);
    for (size_t i = 0; i < n; ++i) {
        const bool outside = (i % 4) == 0;
        for (int k = 0; k < 3; ++k) {
            r[i].o[k] = outside ? rng.uniform(-0.5f * N, 1.5f * N) : rng.uniform(0.0f, N);
            // sums of uniforms: direction components of either sign, some tiny
            r[i].d[k] = rng.uniform(-1, 1) + rng.uniform(-1, 1);
        }
        if (i % 16 == 3) r[i].d[i % 3] = 0.0f; // axis-parallel planes
        r[i].tmin = (i % 8 == 5) ? rng.uniform(0, 40) : 0.0f;
        r[i].tmax = (i % 8 == 6) ? rng.uniform(1, 80) : 1e30f;
    }
    return r;
}

int trace(const char* out)
{
    Built b;
    build_cave(128, 5, 0.5, true, b);
    const LaineKarrasOctreeHip oct(0, b.desc());
    Rng rng{2024};
    Writer w(out);
    auto one_by_one = [&](const std::vector<vx_ray>& rays, auto cast) {
        std::vector<RayHit> h;
        for (const vx_ray& r : rays)
            h.push_back(cast(Vec3{r.o[0], r.o[1], r.o[2]}, Vec3{r.d[0], r.d[1], r.d[2]}, r.tmin, r.tmax));
        w.group(rays, h);
    };
    // 1. castRay, one ray per call
    const std::vector<vx_ray> a = seeded_rays(rng, 3000, 128.0f);
    one_by_one(a, [&](Vec3 o, Vec3 d, float t0, float t1) { return oct.castRay(o, d, t0, t1); });
    // 2. castRayLOD (cone growth 0.01 per unit distance)
    const std::vector<vx_ray> l1 = seeded_rays(rng, 1000, 128.0f);
    one_by_one(l1, [&](Vec3 o, Vec3 d, float t0, float t1) { return oct.castRayLOD(o, d, 0.01f, t0, t1); });
    // 3. castRayScreenSpaceLOD (fovY 1.2 rad, 90 lines: coarse enough to stop)
    const std::vector<vx_ray> l2 = seeded_rays(rng, 1000, 128.0f);
    one_by_one(l2, [&](Vec3 o, Vec3 d, float t0, float t1) { return oct.castRayScreenSpaceLOD(o, d, 1.2f, 90, t0, t1); });
    // 4. castRayWithLOD (fromCamera(1.0, 1080) with bias 10: a cone wide enough to stop rays at bricks)
    const LODParameters lp = LODParameters::fromCamera(1.0f, 1080).withBias(10.0f);
    const std::vector<vx_ray> l3 = seeded_rays(rng, 1000, 128.0f);
    one_by_one(l3, [&](Vec3 o, Vec3 d, float t0, float t1) { return oct.castRayWithLOD(o, d, lp, t0, t1); });
    // 5. castRays: 200,000 rays in chunks, the first 3000 those of group 1
    std::vector<vx_ray> big = a;
    const std::vector<vx_ray> more = seeded_rays(rng, 197000, 128.0f);
    big.insert(big.end(), more.begin(), more.end());
    w.group(big, oct.castRays(big.data(), big.size()));
    // 6. castRays with LOD parameters
    w.group(l3, oct.castRays(l3.data(), l3.size(), &lp));
    return 0;
}

int rate(size_t n)
{
    Built b;
    build_cave(256, 2, 0.5, true, b);
    const LaineKarrasOctreeHip oct(0, b.desc());
    Rng rng{77};
    const std::vector<vx_ray> rays = seeded_rays(rng, n, 256.0f);
    (void)oct.castRays(rays.data(), rays.size()); // warm-up: staging buffers, first launches
    std::vector<double> ms;
    size_t hits = 0;
    for (int k = 0; k < 5; ++k) {
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<RayHit> h = oct.castRays(rays.data(), rays.size());
        ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        hits = 0;
        for (const RayHit& x : h) hits += x.hit ? 1u : 0u;
    }
    std::sort(ms.begin(), ms.end());
    std::printf("{\"rays\": %zu, \"castRays_ms\": %.3f, \"mrays\": %.1f, \"hits\": %zu}\n", n, ms[2],
                (double)n / ms[2] / 1e3, hits);
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "queries" && argc == 3) return queries(argv[2]);
        if (mode == "convert" && argc == 4) return convert(argv[2], argv[3]);
        if (mode == "trace" && argc == 3) return trace(argv[2]);
        if (mode == "rate" && argc == 3) return rate(std::stoul(argv[2]));
        std::fprintf(stderr, "usage: svo_consumer queries OUT | convert IN OUT | trace OUT | rate N\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "svo_consumer: %s\n", e.what());
        return 1;
    }
}
