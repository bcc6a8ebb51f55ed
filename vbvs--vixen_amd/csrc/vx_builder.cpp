// vx_builder.cpp -- host-side scene builder (dense voxel grid -> VIXEN octree
// buffers) and BC1 encoder. Built with -ffp-contract=off so the encoder's float
// arithmetic is reproducible.
//
// Order of the output buffers follows LaineKarrasOctree::rebuild
// (libraries/SVO/src/SVORebuild.cpp:228-732) and VoxelSceneCacher::BuildOctree
// (libraries/CashSystem/src/VoxelSceneCacher.cpp:463-628); the construction is
// this file's own: populated bricks are enumerated directly in Morton order, so
// the breadth-first descriptor order falls out level by level (within a level,
// BFS order from the root equals Morton order of the node cells).
#include "vixen_build.h"
#include "vx_common.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <system_error>
#include <thread>
#include <utility>
#include <vector>

namespace vx {

// Morton interleave of a 21-bit value (libraries/Core/src/MortonEncoding.cpp:18-27).
static inline uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
static inline uint32_t compact3(uint64_t x)
{
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10C30C30C30C30C3ull;
    x = (x ^ (x >> 4)) & 0x100F00F00F00F00Full;
    x = (x ^ (x >> 8)) & 0x1F0000FF0000FFull;
    x = (x ^ (x >> 16)) & 0x1F00000000FFFFull;
    x = (x ^ (x >> 32)) & 0x1FFFFFull;
    return static_cast<uint32_t>(x);
}
static inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

template <class F>
static void parallel_for(uint64_t n, int nthreads, F&& f)
{
    if (nthreads <= 0) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    if (n < 4096 || nthreads == 1) {
        f(uint64_t{0}, n);
        return;
    }
    const uint64_t nt = std::min<uint64_t>(static_cast<uint64_t>(nthreads), n);
    std::vector<std::thread> th;
    th.reserve(nt);
    const uint64_t per = (n + nt - 1) / nt;
    for (uint64_t t = 0; t < nt; ++t) {
        const uint64_t b = std::min(n, t * per), e = std::min(n, (t + 1) * per);
        try {
            th.emplace_back([&f, b, e] { f(b, e); });
        } catch (const std::system_error&) {
            f(b, e); // no thread available: run the range here
        }
    }
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// BC1 encoder: DXT1ColorCompressor::encodeColorHead (DXT1Compressor.cpp:38-55)
// and encodeBlockTyped (:57-111), with validCount 16 and sequential texels.
// ---------------------------------------------------------------------------
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static uint32_t encode_head(const float c0[3], const float c1[3])
{
    uint32_t head = static_cast<uint32_t>(clampi(static_cast<int>(c0[2] * 31.0f + 0.5f), 0, 31));
    head |= static_cast<uint32_t>(
                clampi(static_cast<int>(c0[1] * 63.0f + 0.5f - static_cast<float>(head) * 0.03125f), 0, 63))
            << 5;
    head |= static_cast<uint32_t>(
                clampi(static_cast<int>(c0[0] * 31.0f + 0.5f - static_cast<float>(head) * 0x1p-11f), 0, 31))
            << 11;
    head |= static_cast<uint32_t>(
                clampi(static_cast<int>(c1[2] * 31.0f + 0.5f - static_cast<float>(head) * 0x1p-16f), 0, 31))
            << 16;
    head |= static_cast<uint32_t>(
                clampi(static_cast<int>(c1[1] * 63.0f + 0.5f - static_cast<float>(head) * 0x1p-21f), 0, 63))
            << 21;
    head |= static_cast<uint32_t>(
                clampi(static_cast<int>(c1[0] * 31.0f + 0.5f - static_cast<float>(head) * 0x1p-27f), 0, 31))
            << 27;
    return head;
}

static void decode_head(uint32_t head, float ref[4][3])
{
    const float S = 0x1p-32f;
    ref[0][0] = static_cast<float>(head << 16) * S;
    ref[0][1] = static_cast<float>(head << 21) * S;
    ref[0][2] = static_cast<float>(head << 27) * S;
    ref[1][0] = static_cast<float>(head) * S;
    ref[1][1] = static_cast<float>(head << 5) * S;
    ref[1][2] = static_cast<float>(head << 11) * S;
    const float c23 = 2.0f / 3.0f, c13 = 1.0f / 3.0f;
    for (int i = 0; i < 3; ++i) {
        ref[2][i] = ref[0][i] * c23 + ref[1][i] * c13;
        ref[3][i] = ref[0][i] * c13 + ref[1][i] * c23;
    }
}

static inline float dist2(const float a[3], const float b[3])
{
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

uint64_t bc1_encode16(const float colors[16][3])
{
    int i0 = 0, i1 = 0;
    float best = -FLT_MAX;
    for (int i = 0; i < 16; ++i)
        for (int j = i + 1; j < 16; ++j) {
            const float d = dist2(colors[i], colors[j]);
            if (d > best) {
                i0 = i;
                i1 = j;
                best = d;
            }
        }
    const uint32_t head = encode_head(colors[i0], colors[i1]);
    float ref[4][3];
    decode_head(head, ref);
    uint32_t bits = 0;
    for (int i = 0; i < 16; ++i) {
        int bi = 0;
        float bd = FLT_MAX;
        for (int j = 0; j < 4; ++j) {
            const float d = dist2(colors[i], ref[j]);
            if (d < bd) {
                bi = j;
                bd = d;
            }
        }
        bits |= static_cast<uint32_t>(bi) << (2 * i);
    }
    return static_cast<uint64_t>(head) | (static_cast<uint64_t>(bits) << 32);
}

static inline void rgba_to_float(uint32_t w, float c[3])
{
    c[0] = static_cast<float>(w & 255u) / 255.0f;
    c[1] = static_cast<float>((w >> 8) & 255u) / 255.0f;
    c[2] = static_cast<float>((w >> 16) & 255u) / 255.0f;
}

static void free_built(vx_built_scene* s)
{
    if (!s) return;
    std::free(s->esvo_nodes);
    std::free(s->brick_data);
    std::free(s->bc1_colors);
    std::free(s->brick_coords);
    std::free(s);
}

// The buffers of a scene whose populated brick cells have the Morton codes
// `codes` (ascending, unique): fill(id, code, dst) writes the 512 voxels of
// brick `id` (cell `code`) to dst, z-major (dst[z*64 + y*8 + x]). `who` names
// the entry point in error messages.
template <class Fill>
static vx_status build_tree(const std::vector<uint64_t>& codes, uint32_t N, int with_bc1, int nthreads,
                            Fill&& fill, vx_built_scene** out, const char* who)
{
    const uint32_t B = N / 8;
    int levels = 0; // log2(B): internal levels above the bricks (root included)
    while ((1u << levels) < B) ++levels;
    const uint64_t nb = codes.size();
    if (nb >= 0xFFFFFFull) return fail(VX_ERR_INVALID_ARG, std::string(who) + ": more bricks than 24-bit ids");

    // Internal levels (bottom-up): level l holds the unique codes >> 3l.
    // lvl[0] = bricks, lvl[levels] = {0} (root).
    std::vector<std::vector<uint64_t>> lvl(levels + 1);
    lvl[0] = codes;
    for (int l = 1; l <= levels; ++l) {
        auto& cur = lvl[l];
        for (uint64_t c : lvl[l - 1]) {
            const uint64_t p = c >> 3;
            if (cur.empty() || cur.back() != p) cur.push_back(p);
        }
    }
    if (nb == 0) lvl[levels].assign(1, 0); // an empty scene still has a root

    // BFS descriptor offsets: root first, leaf descriptors last.
    std::vector<uint64_t> offset(levels + 1, 0);
    uint64_t total = 0;
    for (int l = levels; l >= 0; --l) {
        offset[l] = total;
        total += lvl[l].size();
    }
    if (total >= (1ull << 24))
        return fail(VX_ERR_INVALID_ARG, std::string(who) + ": descriptor count exceeds 24-bit far pointers");

    auto* s = static_cast<vx_built_scene*>(std::calloc(1, sizeof(vx_built_scene)));
    if (!s) return fail(VX_ERR_OUT_OF_MEMORY, std::string(who) + ": out of memory");
    s->node_count = total;
    s->brick_count = nb;
    s->internal_count = total - nb;
    s->esvo_nodes = static_cast<uint32_t*>(std::calloc(total * 2, sizeof(uint32_t)));
    s->brick_data = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(1, nb * 512) * sizeof(uint32_t)));
    s->brick_coords = static_cast<uint32_t*>(std::malloc(std::max<uint64_t>(1, nb) * sizeof(uint32_t)));
    if (with_bc1) s->bc1_colors = static_cast<uint64_t*>(std::malloc(std::max<uint64_t>(1, nb * 32) * sizeof(uint64_t)));
    if (!s->esvo_nodes || !s->brick_data || !s->brick_coords || (with_bc1 && !s->bc1_colors)) {
        free_built(s);
        return fail(VX_ERR_OUT_OF_MEMORY, std::string(who) + ": out of memory");
    }

    // Internal descriptors: children of a node are contiguous in the next
    // level down (sorted codes), in octant order.
    for (int l = levels; l >= 1; --l) {
        const auto& cur = lvl[l];
        const auto& child = lvl[l - 1];
        uint64_t ci = 0;
        for (uint64_t j = 0; j < cur.size(); ++j) {
            const uint64_t first = ci;
            uint32_t valid = 0;
            while (ci < child.size() && (child[ci] >> 3) == cur[j]) {
                valid |= 1u << (child[ci] & 7u);
                ++ci;
            }
            const uint64_t ptr = offset[l - 1] + first;
            uint32_t w0 = (valid << 16) | ((l == 1 ? valid : 0u) << 24);
            uint32_t w1 = 0;
            if (valid != 0) {
                if (ptr <= 0x7FFFu) {
                    w0 |= static_cast<uint32_t>(ptr);
                } else {
                    w0 |= 0x8000u;
                    w1 = static_cast<uint32_t>(ptr) & 0xFFFFFFu;
                }
            }
            s->esvo_nodes[2 * (offset[l] + j)] = w0;
            s->esvo_nodes[2 * (offset[l] + j) + 1] = w1;
        }
    }

    // Leaf descriptors, brick pool and BC1 blocks.
    parallel_for(nb, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t id = b; id < e; ++id) {
            const uint64_t m = codes[id];
            const uint32_t bx = compact3(m), by = compact3(m >> 1), bz = compact3(m >> 2);
            s->esvo_nodes[2 * (offset[0] + id)] = 0xFFu << 16 | 0xFFu << 24;
            s->esvo_nodes[2 * (offset[0] + id) + 1] = static_cast<uint32_t>(id) & 0xFFFFFFu;
            s->brick_coords[id] = bx | (by << 10) | (bz << 20);
            uint32_t* dst = s->brick_data + id * 512;
            fill(id, m, dst);
            if (with_bc1) {
                for (int blk = 0; blk < 32; ++blk) {
                    const uint32_t* w = dst + blk * 16;
                    bool any = false;
                    float cols[16][3];
                    for (int t = 0; t < 16; ++t) {
                        if (w[t] != 0u) {
                            any = true;
                            if (with_bc1 == VX_BC1_FROM_MATERIAL) vx_material_id_to_color(w[t], cols[t]);
                            else rgba_to_float(w[t], cols[t]);
                        } else {
                            cols[t][0] = cols[t][1] = cols[t][2] = 0.0f;
                        }
                    }
                    s->bc1_colors[id * 32 + blk] = any ? bc1_encode16(cols) : 0ull;
                }
            }
        }
    });

    // OctreeConfig as VoxelSceneCacher.cpp:592-625 fills it: grid bounds [0,N]^3
    // (voxel units), localToWorld = scale(WORLD_GRID_SIZE = 10), worldToLocal its
    // inverse (the compute shader's 10-unit world; the traversal reads only the
    // grid bounds).
    vx_octree_config& c = s->config;
    std::memset(&c, 0, sizeof c);
    const int L = levels + 3;
    c.esvo_max_scale = 22;
    c.user_max_levels = L;
    c.brick_depth_levels = 3;
    c.brick_size = 8;
    c.min_esvo_scale = 22 - L + 1;
    c.brick_esvo_scale = 22 - (L - 1 - (L - 3));
    c.bricks_per_axis = static_cast<int32_t>(B);
    const float Nf = static_cast<float>(N);
    for (int i = 0; i < 3; ++i) {
        c.grid_min[i] = 0.0f;
        c.grid_max[i] = Nf;
    }
    for (int i = 0; i < 16; ++i) c.local_to_world[i] = c.world_to_local[i] = 0.0f;
    for (int i = 0; i < 3; ++i) {
        c.local_to_world[i * 5] = 10.0f;
        c.world_to_local[i * 5] = 0.1f;
    }
    c.local_to_world[15] = c.world_to_local[15] = 1.0f;
    *out = s;
    return VX_OK;
}

static bool valid_edge(uint32_t N) { return N >= 16 && N <= 4096 && (N & (N - 1)) == 0; }

static vx_status build_from_dense(const uint32_t* grid, uint32_t N, int with_bc1, int nthreads,
                                  vx_built_scene** out)
{
    if (!grid || !out) return fail(VX_ERR_INVALID_ARG, "vx_build_from_dense: null argument");
    if (with_bc1 < VX_BC1_NONE || with_bc1 > VX_BC1_FROM_MATERIAL)
        return fail(VX_ERR_INVALID_ARG, "vx_build_from_dense: with_bc1 must be VX_BC1_NONE, _FROM_RGBA8 or _FROM_MATERIAL");
    if (!valid_edge(N))
        return fail(VX_ERR_INVALID_ARG, "vx_build_from_dense: N must be a power of two in [16, 4096]");
    *out = nullptr;
    const uint32_t B = N / 8;
    const uint64_t cells = static_cast<uint64_t>(B) * B * B;

    // 1. Occupancy of every brick cell, visited in Morton order.
    std::vector<uint8_t> occ(cells, 0);
    parallel_for(cells, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t m = b; m < e; ++m) {
            const uint32_t bx = compact3(m), by = compact3(m >> 1), bz = compact3(m >> 2);
            bool any = false;
            for (uint32_t z = 0; z < 8 && !any; ++z)
                for (uint32_t y = 0; y < 8 && !any; ++y) {
                    const uint32_t* row = grid + (static_cast<uint64_t>(bz * 8 + z) * N + (by * 8 + y)) * N + bx * 8;
                    for (uint32_t x = 0; x < 8; ++x)
                        if (row[x] != 0u) {
                            any = true;
                            break;
                        }
                }
            occ[m] = any ? 1 : 0;
        }
    });

    // 2. Brick ids = Morton ranks of the populated cells.
    std::vector<uint64_t> codes;
    for (uint64_t m = 0; m < cells; ++m)
        if (occ[m]) codes.push_back(m);
    auto fill = [&](uint64_t, uint64_t m, uint32_t* dst) {
        const uint32_t bx = compact3(m), by = compact3(m >> 1), bz = compact3(m >> 2);
        for (uint32_t z = 0; z < 8; ++z)
            for (uint32_t y = 0; y < 8; ++y) {
                const uint32_t* row = grid + (static_cast<uint64_t>(bz * 8 + z) * N + (by * 8 + y)) * N + bx * 8;
                std::memcpy(dst + z * 64 + y * 8, row, 8 * sizeof(uint32_t));
            }
    };
    return build_tree(codes, N, with_bc1, nthreads, fill, out, "vx_build_from_dense");
}

static vx_status build_from_bricks(const uint32_t* cells, const uint32_t* voxels, uint64_t n, uint32_t N,
                                   int with_bc1, int nthreads, vx_built_scene** out)
{
    if (!out || (n && (!cells || !voxels))) return fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: null argument");
    if (with_bc1 < VX_BC1_NONE || with_bc1 > VX_BC1_FROM_MATERIAL)
        return fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: with_bc1 must be VX_BC1_NONE, _FROM_RGBA8 or _FROM_MATERIAL");
    if (!valid_edge(N))
        return fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: N must be a power of two in [16, 4096]");
    *out = nullptr;
    const uint32_t B = N / 8;
    // (Morton code, input index) of every brick with a solid voxel
    std::vector<std::pair<uint64_t, uint64_t>> order;
    order.reserve(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t x = cells[3 * i], y = cells[3 * i + 1], z = cells[3 * i + 2];
        if (x >= B || y >= B || z >= B)
            return fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: brick cell " + std::to_string(i) +
                                                " outside the " + std::to_string(B) + "^3 brick grid");
        const uint32_t* v = voxels + i * 512;
        if (std::any_of(v, v + 512, [](uint32_t w) { return w != 0u; })) order.emplace_back(morton3(x, y, z), i);
    }
    std::sort(order.begin(), order.end());
    for (size_t k = 1; k < order.size(); ++k)
        if (order[k].first == order[k - 1].first)
            return fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: bricks " + std::to_string(order[k - 1].second) +
                                                " and " + std::to_string(order[k].second) + " share a cell");
    std::vector<uint64_t> codes(order.size());
    for (size_t k = 0; k < order.size(); ++k) codes[k] = order[k].first;
    auto fill = [&](uint64_t id, uint64_t, uint32_t* dst) {
        std::memcpy(dst, voxels + order[id].second * 512, 512 * sizeof(uint32_t));
    };
    return build_tree(codes, N, with_bc1, nthreads, fill, out, "vx_build_from_bricks");
}

} // namespace vx

extern "C" {

vx_status vx_build_from_dense(const uint32_t* grid, uint32_t N, int with_bc1, int nthreads, vx_built_scene** out)
{
    try {
        return vx::build_from_dense(grid, N, with_bc1, nthreads, out);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_build_from_dense: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_ARG, "vx_build_from_dense: unexpected exception");
    }
}

vx_status vx_build_from_bricks(const uint32_t* brick_cells, const uint32_t* brick_voxels, uint64_t n, uint32_t N,
                               int with_bc1, int nthreads, vx_built_scene** out)
{
    try {
        return vx::build_from_bricks(brick_cells, brick_voxels, n, N, with_bc1, nthreads, out);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_build_from_bricks: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_ARG, "vx_build_from_bricks: unexpected exception");
    }
}

void vx_built_scene_free(vx_built_scene* s) { vx::free_built(s); }

uint64_t vx_bc1_encode_block(const float colors[16][3]) { return vx::bc1_encode16(colors); }

} // extern "C"
