// vx_scenes.cpp -- seeded synthetic scenes for the parity tests and the
// benchmark: a Cornell-like room, an fBm value-noise cave and an urban box grid.
// Scene space is voxel units, world = [0,N]^3, y up. All randomness comes from
// SplitMix64 streams keyed by (seed, purpose, coordinates), so a scene depends
// only on its arguments.
#include "vixen_build.h"
#include "vx_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

namespace vx {

static inline uint64_t splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c)
{
    return splitmix(splitmix(splitmix(seed ^ a * 0xD6E8FEB86659FD93ull) ^ b * 0xA0761D6478BD642Full) ^
                    c * 0xE7037ED1A0B428DBull);
}
static inline double unit(uint64_t h) { return static_cast<double>(h >> 11) * 0x1p-53; }

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(splitmix(seed ^ 0x5EED5EED5EEDull)) {}
    uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return unit(next()); }
    int range(int lo, int hi) { return lo + static_cast<int>(next() % static_cast<uint64_t>(hi - lo + 1)); }
};

static inline uint32_t rgba(int r, int g, int b)
{
    auto c = [](int v) { return static_cast<uint32_t>(v < 0 ? 0 : (v > 255 ? 255 : v)); };
    return c(r) | (c(g) << 8) | (c(b) << 16) | 0xFF000000u;
}

template <class F>
static void par_z(uint32_t N, F&& f)
{
    const unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        auto run = [&, t] {
            for (uint32_t z = t; z < N; z += nt) f(z);
        };
        try {
            th.emplace_back(run);
        } catch (const std::system_error&) {
            run(); // no thread available: run the slices here
        }
    }
    for (auto& x : th) x.join();
}

// Camera basis from a look direction: right = dir x worldUp, up = right x dir.
static void make_camera(const double pos[3], const double dir_in[3], double fov_deg, double aspect, vx_camera* cam)
{
    double d[3] = {dir_in[0], dir_in[1], dir_in[2]};
    double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double& v : d) v /= n;
    double wu[3] = {0.0, 1.0, 0.0};
    if (std::fabs(d[1]) > 0.99) {
        wu[1] = 0.0;
        wu[2] = 1.0;
    }
    double r[3] = {d[1] * wu[2] - d[2] * wu[1], d[2] * wu[0] - d[0] * wu[2], d[0] * wu[1] - d[1] * wu[0]};
    n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& v : r) v /= n;
    const double u[3] = {r[1] * d[2] - r[2] * d[1], r[2] * d[0] - r[0] * d[2], r[0] * d[1] - r[1] * d[0]};
    std::memset(cam, 0, sizeof *cam);
    for (int i = 0; i < 3; ++i) {
        cam->pos[i] = static_cast<float>(pos[i]);
        cam->dir[i] = static_cast<float>(d[i]);
        cam->up[i] = static_cast<float>(u[i]);
        cam->right[i] = static_cast<float>(r[i]);
    }
    cam->fov_deg = static_cast<float>(fov_deg);
    cam->aspect = static_cast<float>(aspect);
}

static bool pow2(uint32_t N) { return N >= 16 && N <= 4096 && (N & (N - 1)) == 0; }

// ---------------------------------------------------------------------------
// Cornell-like room
// ---------------------------------------------------------------------------
static vx_status gen_cornell(uint32_t N, uint64_t seed, uint32_t* grid, vx_camera* cam)
{
    if (!grid || !cam || !pow2(N) || N < 32) return fail(VX_ERR_INVALID_ARG, "vx_gen_cornell: bad arguments");
    Rng rng(seed);
    const int n = static_cast<int>(N);
    // two boxes on the floor at seeded positions, kept off the walls
    const int b1x = rng.range(2, n / 2 - 14), b1z = rng.range(n / 3, n - 16);
    const int b2x = rng.range(n / 2 + 1, n - 15), b2z = rng.range(n / 4, n - 16);
    par_z(N, [&](uint32_t z) {
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const int zz = static_cast<int>(z);
                int r = -1, g = 0, b = 0;
                if (x == 0) { r = 255; g = 0; b = 0; }
                else if (x == n - 1) { r = 0; g = 255; b = 0; }
                else if (y == 0 || y == n - 1 || zz == n - 1) { r = g = b = 200; }
                else if (x >= b1x && x < b1x + 12 && y >= 1 && y < 13 && zz >= b1z && zz < b1z + 12) { r = g = b = 200; }
                else if (x >= b2x && x < b2x + 12 && y >= 1 && y < 25 && zz >= b2z && zz < b2z + 12) { r = g = b = 200; }
                uint32_t w = 0;
                if (r >= 0) {
                    const uint64_t h = hash4(seed, 0xC0, static_cast<uint64_t>(x) | (static_cast<uint64_t>(y) << 16),
                                             static_cast<uint64_t>(zz));
                    const int nr = static_cast<int>(h % 17) - 8, ng = static_cast<int>((h >> 8) % 17) - 8,
                              nbv = static_cast<int>((h >> 16) % 17) - 8;
                    w = rgba(r + nr, g + ng, b + nbv);
                }
                grid[(static_cast<uint64_t>(z) * N + y) * N + x] = w;
            }
    });
    const double pos[3] = {0.5 * n, 0.5 * n, 0.0}, dir[3] = {0.0, 0.0, 1.0};
    make_camera(pos, dir, 60.0, 1.0, cam);
    return VX_OK;
}

// ---------------------------------------------------------------------------
// fBm value-noise cave
// ---------------------------------------------------------------------------
struct Lattice {
    int n;                   // points per axis
    std::vector<float> v;    // n^3 values in [0,1)
};

static double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

static vx_status gen_cave(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam)
{
    if (!grid || !cam || !pow2(N) || !(occupancy > 0.0 && occupancy < 1.0))
        return fail(VX_ERR_INVALID_ARG, "vx_gen_cave: bad arguments");
    const int octaves = 4;
    const double base_freq = 1.0 / 32.0, lacunarity = 2.0, gain = 0.5;
    Lattice lat[octaves];
    for (int o = 0; o < octaves; ++o) {
        const double f = base_freq * std::pow(lacunarity, o);
        lat[o].n = static_cast<int>(std::ceil(N * f)) + 2;
        const int ln = lat[o].n;
        lat[o].v.resize(static_cast<size_t>(ln) * ln * ln);
        for (int z = 0; z < ln; ++z)
            for (int y = 0; y < ln; ++y)
                for (int x = 0; x < ln; ++x)
                    lat[o].v[(static_cast<size_t>(z) * ln + y) * ln + x] = static_cast<float>(
                        unit(hash4(seed, 0xFB00 + o, static_cast<uint64_t>(x) | (static_cast<uint64_t>(y) << 20),
                                   static_cast<uint64_t>(z))));
    }
    const uint64_t total = static_cast<uint64_t>(N) * N * N;
    std::vector<float> val(total);
    par_z(N, [&](uint32_t z) {
        for (uint32_t y = 0; y < N; ++y)
            for (uint32_t x = 0; x < N; ++x) {
                double sum = 0.0, amp = 1.0, f = base_freq;
                for (int o = 0; o < octaves; ++o) {
                    const double px = (x + 0.5) * f, py = (y + 0.5) * f, pz = (z + 0.5) * f;
                    const int ix = static_cast<int>(px), iy = static_cast<int>(py), iz = static_cast<int>(pz);
                    const double fx = smooth(px - ix), fy = smooth(py - iy), fz = smooth(pz - iz);
                    const int ln = lat[o].n;
                    auto L = [&](int a, int b, int c) {
                        return static_cast<double>(lat[o].v[(static_cast<size_t>(c) * ln + b) * ln + a]);
                    };
                    const double c00 = L(ix, iy, iz) + (L(ix + 1, iy, iz) - L(ix, iy, iz)) * fx;
                    const double c10 = L(ix, iy + 1, iz) + (L(ix + 1, iy + 1, iz) - L(ix, iy + 1, iz)) * fx;
                    const double c01 = L(ix, iy, iz + 1) + (L(ix + 1, iy, iz + 1) - L(ix, iy, iz + 1)) * fx;
                    const double c11 = L(ix, iy + 1, iz + 1) + (L(ix + 1, iy + 1, iz + 1) - L(ix, iy + 1, iz + 1)) * fx;
                    const double c0 = c00 + (c10 - c00) * fy, c1 = c01 + (c11 - c01) * fy;
                    sum += amp * (c0 + (c1 - c0) * fz);
                    amp *= gain;
                    f *= lacunarity;
                }
                val[(static_cast<uint64_t>(z) * N + y) * N + x] = static_cast<float>(sum);
            }
    });
    // threshold: the (1 - occupancy) quantile, so that `occupancy` of the voxels exceed it
    std::vector<float> tmp(val);
    const uint64_t k = static_cast<uint64_t>((1.0 - occupancy) * static_cast<double>(total));
    std::nth_element(tmp.begin(), tmp.begin() + static_cast<std::ptrdiff_t>(std::min(k, total - 1)), tmp.end());
    const float thr = tmp[std::min(k, total - 1)];
    std::vector<float>().swap(tmp);
    par_z(N, [&](uint32_t z) {
        for (uint32_t y = 0; y < N; ++y)
            for (uint32_t x = 0; x < N; ++x) {
                const uint64_t i = (static_cast<uint64_t>(z) * N + y) * N + x;
                uint32_t w = 0;
                if (val[i] > thr) {
                    const double hgt = static_cast<double>(y) / static_cast<double>(N - 1);
                    const uint64_t h = hash4(seed, 0xCA7E, static_cast<uint64_t>(x) | (static_cast<uint64_t>(y) << 20),
                                             static_cast<uint64_t>(z));
                    const int nr = static_cast<int>(h % 25) - 12, ng = static_cast<int>((h >> 8) % 25) - 12,
                              nbv = static_cast<int>((h >> 16) % 25) - 12;
                    w = rgba(static_cast<int>(90 + 110 * hgt) + nr, static_cast<int>(60 + 110 * hgt) + ng,
                             static_cast<int>(40 + 100 * hgt) + nbv);
                }
                grid[i] = w;
            }
    });
    // camera: a seeded empty cell with clearance, seeded look direction
    Rng rng(seed ^ 0xCA3E7A);
    auto empty_box = [&](int cx, int cy, int cz, int r) {
        for (int z = cz - r; z <= cz + r; ++z)
            for (int y = cy - r; y <= cy + r; ++y)
                for (int x = cx - r; x <= cx + r; ++x) {
                    if (x < 0 || y < 0 || z < 0 || x >= static_cast<int>(N) || y >= static_cast<int>(N) ||
                        z >= static_cast<int>(N))
                        return false;
                    if (grid[(static_cast<uint64_t>(z) * N + y) * N + x] != 0u) return false;
                }
        return true;
    };
    int cx = static_cast<int>(N / 2), cy = static_cast<int>(N / 2), cz = static_cast<int>(N / 2);
    bool found = false;
    for (int r = 3; r >= 0 && !found; --r)
        for (int tries = 0; tries < 200000 && !found; ++tries) {
            const int x = rng.range(0, static_cast<int>(N) - 1), y = rng.range(0, static_cast<int>(N) - 1),
                      z = rng.range(0, static_cast<int>(N) - 1);
            if (empty_box(x, y, z, r)) {
                cx = x;
                cy = y;
                cz = z;
                found = true;
            }
        }
    double dir[3];
    {
        const double u = 2.0 * rng.uniform() - 1.0, phi = 2.0 * M_PI * rng.uniform();
        const double s = std::sqrt(std::max(0.0, 1.0 - u * u));
        dir[0] = s * std::cos(phi);
        dir[1] = u;
        dir[2] = s * std::sin(phi);
    }
    const double pos[3] = {cx + 0.5, cy + 0.5, cz + 0.5};
    make_camera(pos, dir, 70.0, 16.0 / 9.0, cam);
    return VX_OK;
}

// ---------------------------------------------------------------------------
// Urban box grid
// ---------------------------------------------------------------------------
static vx_status gen_urban(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam)
{
    if (!grid || !cam || !pow2(N) || N < 64 || !(occupancy > 0.0 && occupancy <= 1.0))
        return fail(VX_ERR_INVALID_ARG, "vx_gen_urban: bad arguments");
    const int pitch = 32, lots = static_cast<int>(N) / pitch;
    struct Lot {
        int x0, x1, z0, z1, h;
        uint32_t r, g, b;
        bool built;
    };
    std::vector<Lot> lot(static_cast<size_t>(lots) * lots);
    Rng rng(seed ^ 0x0C17);
    const double hs = static_cast<double>(N) / 1024.0;
    for (int j = 0; j < lots; ++j)
        for (int i = 0; i < lots; ++i) {
            Lot& L = lot[static_cast<size_t>(j) * lots + i];
            const int fx = rng.range(24, 30), fz = rng.range(24, 30);
            const int ox = rng.range(1, pitch - 1 - fx), oz = rng.range(1, pitch - 1 - fz);
            L.x0 = i * pitch + ox;
            L.x1 = L.x0 + fx;
            L.z0 = j * pitch + oz;
            L.z1 = L.z0 + fz;
            L.h = std::max(2, static_cast<int>((64.0 + 836.0 * rng.uniform()) * hs));
            L.r = static_cast<uint32_t>(rng.range(60, 230));
            L.g = static_cast<uint32_t>(rng.range(60, 230));
            L.b = static_cast<uint32_t>(rng.range(60, 230));
            L.built = rng.uniform() < occupancy;
        }
    par_z(N, [&](uint32_t z) {
        const int zz = static_cast<int>(z), lj = zz / pitch;
        for (uint32_t y = 0; y < N; ++y)
            for (uint32_t x = 0; x < N; ++x) {
                const int xx = static_cast<int>(x), li = xx / pitch;
                uint32_t w = 0;
                const Lot& L = lot[static_cast<size_t>(lj) * lots + li];
                const uint64_t key = static_cast<uint64_t>(x) | (static_cast<uint64_t>(y) << 20);
                if (L.built && xx >= L.x0 && xx < L.x1 && zz >= L.z0 && zz < L.z1 && static_cast<int>(y) < L.h) {
                    const uint64_t h = hash4(seed, 0xB1D, key, z);
                    const int n = static_cast<int>(h % 13) - 6;
                    w = rgba(static_cast<int>(L.r) + n, static_cast<int>(L.g) + n, static_cast<int>(L.b) + n);
                } else if (y == 0) {
                    const uint64_t h = hash4(seed, 0x6D, key, z);
                    const int n = static_cast<int>(h % 9) - 4;
                    w = rgba(90 + n, 90 + n, 95 + n);
                }
                grid[(static_cast<uint64_t>(z) * N + y) * N + x] = w;
            }
    });
    const double pos[3] = {(0.1 + 0.8 * rng.uniform()) * N, 0.97 * N, (0.1 + 0.2 * rng.uniform()) * N};
    const double tgt[3] = {(0.3 + 0.4 * rng.uniform()) * N, 0.0, (0.5 + 0.3 * rng.uniform()) * N};
    const double dir[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
    make_camera(pos, dir, 70.0, 16.0 / 9.0, cam);
    return VX_OK;
}

} // namespace vx

extern "C" {

vx_status vx_gen_cornell(uint32_t N, uint64_t seed, uint32_t* grid, vx_camera* cam)
{
    try {
        return vx::gen_cornell(N, seed, grid, cam);
    } catch (...) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_gen_cornell: exception");
    }
}

vx_status vx_gen_cave(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam)
{
    try {
        return vx::gen_cave(N, seed, occupancy, grid, cam);
    } catch (...) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_gen_cave: exception");
    }
}

vx_status vx_gen_urban(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam)
{
    try {
        return vx::gen_urban(N, seed, occupancy, grid, cam);
    } catch (...) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_gen_urban: exception");
    }
}

} // extern "C"
