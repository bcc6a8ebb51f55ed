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
//   deterministic  (row, pixel) pairs are sorted by row with a stable LSD
//                  radix sort written here (equal rows keep pixel order,
//                  misses last); a run head sums its run in a fixed order (one
//                  thread in sequence for runs of <= 32 pairs, one wave for
//                  runs of <= 4096, else one workgroup: lane or thread t takes
//                  elements t, t + width, ..., then a fixed tree), so the
//                  per-row sums, added without conflicts,
//                  depend only on the input: identical bits on every run. (A
//                  decoupled look-back reduce-by-key is not used: the order in
//                  which it combines the partial sums of a run spanning several
//                  tiles depends on timing.)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>

#include "vx_internal.hpp"

namespace vx {

constexpr uint32_t kAlbBlock = 256;

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
// The brick's 16 mask words (one 64-byte line) are read with four independent
// 16-byte loads and counted with per-word masks, rather than a loop of dependent
// loads up to the voxel's word.
__device__ __forceinline__ uint32_t voxel_row(const SceneView& s, const uint32_t* __restrict__ base, uint32_t brick,
                                              uint32_t lin)
{
    const uint4* m = reinterpret_cast<const uint4*>(s.occ + (size_t)brick * 16u);
    const uint4 q0 = m[0], q1 = m[1], q2 = m[2], q3 = m[3];
    const uint32_t words[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                                q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
    const uint32_t wq = lin >> 5, low = (1u << (lin & 31u)) - 1u;
    uint32_t r = base[brick];
#pragma unroll
    for (uint32_t w = 0; w < 16u; ++w) r += __builtin_popcount(words[w] & (w < wq ? ~0u : (w == wq ? low : 0u)));
    return r;
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

// Pixel of lane threadIdx.x in 16x16 tile `tile` (row-major tiles, 8x8 per wave).
__device__ __forceinline__ bool tile_pixel(uint32_t w, uint32_t h, uint32_t tile, uint32_t& px, uint32_t& py)
{
    const uint32_t tiles_x = (w + 15u) / 16u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
    return tile < tiles_x * ((h + 15u) / 16u) && px < w && py < h;
}

__global__ void __launch_bounds__(kAlbBlock)
    k_albedo_forward(SceneView s, AlbCam c, uint32_t w, uint32_t h, uint32_t per_xcd, const uint32_t* __restrict__ base,
                     const float* __restrict__ albedo, float* __restrict__ image, uint32_t* __restrict__ hit_row,
                     float* __restrict__ hit_shade)
{
    uint32_t px, py;
    if (!tile_pixel(w, h, xcd_stripe_chunk<kRenderStripe>(per_xcd), px, py)) return;
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
// cross-lane operations see the whole wave. n_voxels is the scene's voxel
// count: a stored row at or above it (the miss marker, or anything else the
// caller's buffer holds there) counts as a miss, so no index leaves the
// gradient; misses get the sort key n_voxels, which sorts after every row.
template <bool RECOMPUTE, bool ATOMIC>
__global__ void __launch_bounds__(kAlbBlock)
    k_albedo_adjoint(SceneView s, AlbCam c, uint32_t w, uint32_t h, uint32_t per_xcd, const uint32_t* __restrict__ base,
                     const float* __restrict__ grad_image, const uint32_t* __restrict__ hit_row,
                     const float* __restrict__ hit_shade, float* __restrict__ grad, uint32_t* __restrict__ keys,
                     Rgb* __restrict__ vals, uint32_t* __restrict__ pix, uint32_t n_voxels)
{
    uint32_t px, py;
    const bool inside = tile_pixel(w, h, xcd_stripe_chunk<kRenderStripe>(per_xcd), px, py);
    const uint64_t p = (uint64_t)py * w + px;
    uint32_t row = kNoVoxel;
    float shade = 0.0f;
    if (inside) {
        if (RECOMPUTE) {
            row = trace_pixel(s, c, base, px, py, w, h, shade);
        } else {
            row = hit_row[p];
            if (row >= n_voxels) row = kNoVoxel;
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
        keys[p] = row == kNoVoxel ? n_voxels : row; // misses sort last
        vals[p] = Rgb{gr, gg, gb};
        pix[p] = (uint32_t)p;
    }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs, 8 key bits per pass. A pass is
// three launches: k_sort_count (each tile of kSortTile pairs counts its
// digits), k_sort_scan (one workgroup per digit: exclusive scan of that digit's
// counts over the tiles, and the digit's total) and k_sort_scatter (each tile
// writes its pairs to digit base + earlier tiles' count + rank in the tile).
// A tile is kSortWaves contiguous slices of kSortRounds x 64 pairs, one per
// wave; a wave ranks its slice round by round, matching digits across its lanes
// (one ballot per key bit) and keeping its own digit counters in LDS, so pairs
// are ranked in index order without atomics: the sort is stable, and every
// count is an integer, so the output order does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortRounds = 8; // pairs per thread and tile
constexpr uint32_t kSortTile = kSortThreads * kSortRounds;
constexpr uint32_t kRadix = 256;
static_assert(kSortThreads == kRadix, "one thread per digit in the scan and bookkeeping steps");

// Lanes among `active` whose 8-bit digit equals this lane's (every lane of the
// wave executes it).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active)
{
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t ones = __ballot(bit);
        m &= bit ? ones : ~ones;
    }
    return m;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m, uint32_t lane)
{
    return (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Exclusive scan of one value per thread over a workgroup of kSortThreads;
// `total` receives the sum of all values. Integer adds: the order is irrelevant.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total)
{
    __shared__ uint32_t wsum[kSortWaves];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63u) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; ++w) {
        before += w < wave ? wsum[w] : 0u;
        all += wsum[w];
    }
    __syncthreads(); // wsum is reused by the next call
    total = all;
    return before + x - v;
}

// Digits of the wave's slice of the tile, counted into its LDS counters c[]
// (zeroed by the caller); rank[k] receives each pair's rank among the slice's
// pairs of its digit. The wave's LDS reads and writes execute in order, so a
// round's reads see the previous round's counts.
__device__ __forceinline__ void rank_slice(const uint32_t (&d)[kSortRounds], uint32_t base, uint32_t n,
                                           uint32_t* c, uint32_t (&rank)[kSortRounds])
{
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t k = 0; k < kSortRounds; ++k) {
        const bool act = base + k * 64u + lane < n;
        const uint64_t peers = match_digit(d[k], __ballot(act));
        const uint32_t below = lanes_below(peers, lane);
        const uint32_t before = act ? c[d[k]] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (act && below == 0u) c[d[k]] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[k] = before + below;
    }
}

__global__ void __launch_bounds__(kSortThreads) k_sort_count(const uint32_t* __restrict__ keys, uint32_t n,
                                                             uint32_t shift, uint32_t n_tiles,
                                                             uint32_t* __restrict__ count)
{
    __shared__ uint32_t c[kSortWaves][kRadix];
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; ++w) c[w][t] = 0u;
    const uint32_t base = blockIdx.x * kSortTile + wave * (kSortRounds * 64u); // this wave's slice
    uint32_t d[kSortRounds], rank[kSortRounds];
#pragma unroll
    for (uint32_t k = 0; k < kSortRounds; ++k) { // every load issued before any is used
        const uint32_t i = base + k * 64u + lane;
        d[k] = i < n ? (keys[i] >> shift) & 0xFFu : 0u;
    }
    __syncthreads();
    rank_slice(d, base, n, c[wave], rank);
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; ++w) sum += c[w][t];
    count[t * n_tiles + blockIdx.x] = sum;
}

// Workgroup d: exclusive scan of count[d][0 .. n_tiles) in place, total[d] = the
// digit's count.
__global__ void __launch_bounds__(kSortThreads) k_sort_scan(uint32_t* __restrict__ count, uint32_t n_tiles,
                                                            uint32_t* __restrict__ total)
{
    uint32_t* c = count + (size_t)blockIdx.x * n_tiles;
    uint32_t running = 0;
    for (uint32_t b = 0; b < n_tiles; b += kSortThreads) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < n_tiles ? c[i] : 0u;
        uint32_t sum;
        const uint32_t ex = block_exclusive_scan(v, sum);
        if (i < n_tiles) c[i] = running + ex;
        running += sum;
    }
    if (threadIdx.x == 0u) total[blockIdx.x] = running;
}

// GATHER (the last pass): the values are pixel indices, and each pair's
// contribution contrib[pixel] is also written to sorted[pos], so the run sums
// read contributions in sorted order instead of gathering them one by one.
template <bool GATHER>
__global__ void __launch_bounds__(kSortThreads)
    k_sort_scatter(const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t n,
                   uint32_t shift, uint32_t n_tiles, const uint32_t* __restrict__ count,
                   const uint32_t* __restrict__ total, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                   const Rgb* __restrict__ contrib, Rgb* __restrict__ sorted)
{
    __shared__ uint32_t c[kSortWaves][kRadix]; // per-wave digit counts, then each wave's first slot per digit
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; ++w) c[w][t] = 0u;
    const uint32_t base = blockIdx.x * kSortTile + wave * (kSortRounds * 64u);
    uint32_t key[kSortRounds], val[kSortRounds], d[kSortRounds], rank[kSortRounds];
#pragma unroll
    for (uint32_t k = 0; k < kSortRounds; ++k) {
        const uint32_t i = base + k * 64u + lane;
        key[k] = i < n ? keys_in[i] : 0u;
        val[k] = i < n ? vals_in[i] : 0u;
    }
    Rgb g[GATHER ? kSortRounds : 1];
    if (GATHER) {
#pragma unroll
        for (uint32_t k = 0; k < kSortRounds; ++k) // in flight while the pairs are ranked
            if (base + k * 64u + lane < n) g[k] = contrib[val[k]];
    }
    uint32_t all;
    const uint32_t tile_base = block_exclusive_scan(total[t], all) + count[t * n_tiles + blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < kSortRounds; ++k) d[k] = (key[k] >> shift) & 0xFFu;
    rank_slice(d, base, n, c[wave], rank);
    __syncthreads();
    // digit t: wave w's pairs start after the tile's earlier waves' pairs of that digit
    uint32_t first = tile_base;
#pragma unroll
    for (uint32_t w = 0; w < kSortWaves; ++w) {
        const uint32_t cw = c[w][t];
        c[w][t] = first;
        first += cw;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kSortRounds; ++k) {
        if (base + k * 64u + lane < n) {
            const uint32_t pos = c[wave][d[k]] + rank[k];
            keys_out[pos] = key[k];
            vals_out[pos] = val[k];
            if (GATHER) sorted[pos] = g[k];
        }
    }
}

// Fixed-order run sums over the sorted pairs. Every run's sum is a function of
// its elements and length only, so the gradient bits do not depend on
// scheduling. Each workgroup takes a window of kRunWindow positions, kRunItems
// consecutive ones per thread: it marks the run starts (key differs from the
// previous one) in an LDS bit mask and finds each run's end at the next start;
// a run still open at the window's end is either closed within the next
// kShortRun positions (independent loads) or certainly long. The runs of at
// most kShortRun pairs are compacted into an LDS list and added in order, one
// thread per run. Longer runs (voxels that cover many pixels) go to a second
// LDS list of at most kLongPerWindow runs, with their end, or with end 0 when it
// lies past the window, and the same workgroup then sums them (sum_long_runs).
// Misses (key miss_key) sort last and are skipped. Rows are unique across runs:
// no conflicts.
constexpr uint32_t kShortRun = 32;
constexpr uint32_t kRunItems = 2;
constexpr uint32_t kRunWindow = kAlbBlock * kRunItems;
static_assert(kAlbBlock == kSortThreads, "k_reduce_runs uses block_exclusive_scan");
static_assert(32u % kRunItems == 0u && kRunWindow / 32u <= kAlbBlock, "a thread's run starts fit in one mask word");
// long runs start at least kShortRun + 1 positions apart
constexpr uint32_t kLongPerWindow = kRunWindow / (kShortRun + 1u) + 1u;

// Runs longer than this are summed by a whole workgroup, shorter long runs by one wave.
constexpr uint32_t kWaveRun = 4096;

// Sum of a run's contributions [b, e) by one wave in a fixed order: lane l adds
// elements b+l, b+l+128, ... into one partial sum and b+l+64, b+l+192, ... into
// another, loading four steps ahead; then s0 + s1 per lane and a halving tree
// over the lanes (h = 32, 16, ..., 1: lane t < h adds lane t + h's value).
// Every lane of the wave calls it; lane 0's result is the sum.
__device__ __forceinline__ void wave_run_sum(const Rgb* __restrict__ vals_sorted, uint32_t b, uint32_t e,
                                             uint32_t lane, float out[3])
{
    float s0[3] = {0.0f, 0.0f, 0.0f}, s1[3] = {0.0f, 0.0f, 0.0f};
    uint32_t i = b + lane;
    for (; i + 7u * 64u < e; i += 8u * 64u) { // four steps: eight loads in flight
        Rgb v[8];
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) v[k] = vals_sorted[i + k * 64u];
#pragma unroll
        for (uint32_t k = 0; k < 8u; k += 2u) {
            s0[0] += v[k].r; s0[1] += v[k].g; s0[2] += v[k].b;
            s1[0] += v[k + 1].r; s1[1] += v[k + 1].g; s1[2] += v[k + 1].b;
        }
    }
    for (; i + 64u < e; i += 128u) {
        const Rgb v0 = vals_sorted[i], v1 = vals_sorted[i + 64u];
        s0[0] += v0.r; s0[1] += v0.g; s0[2] += v0.b;
        s1[0] += v1.r; s1[1] += v1.g; s1[2] += v1.b;
    }
    if (i < e) {
        const Rgb v0 = vals_sorted[i];
        s0[0] += v0.r; s0[1] += v0.g; s0[2] += v0.b;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float x = s0[c] + s1[c];
#pragma unroll
        for (uint32_t h = 32u; h > 0u; h >>= 1) x += __shfl_down(x, h, 64); // lanes >= h: unused
        out[c] = x;
    }
}

// End of the run of `row` that starts at b and continues past its window: the
// first position after b whose key differs, or n. Bisection by one wave with
// one probe per lane (keys are sorted: the probes that still hold the row are a
// prefix). Every lane of the wave calls it.
__device__ __forceinline__ uint32_t wave_run_end(const uint32_t* __restrict__ keys_sorted, uint32_t n, uint32_t b,
                                                 uint32_t row, uint32_t lane)
{
    // keys[lo - 1] == row; hi == n or keys[hi] != row; the end is in [lo, hi]
    uint32_t lo = b + 1u, hi = n;
    while (lo < hi) {
        const uint32_t q = lo + (uint32_t)(((uint64_t)(hi - lo) * lane) / 64u);
        const uint32_t held = (uint32_t)__popcll(__ballot(keys_sorted[q] == row)); // probes 0 .. held-1
        const uint32_t nlo = held > 0u ? (uint32_t)__shfl((int)q, (int)held - 1, 64) + 1u : lo;
        const uint32_t nhi = held < 64u ? (uint32_t)__shfl((int)q, (int)held, 64) : hi;
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

// A window's long runs [b, e), listed in LDS by k_reduce_runs (an end listed as
// 0 lies past the window and is found by wave_run_end). Wave w takes runs w,
// w+4, ...: a run of at most kWaveRun pairs is summed by that wave
// (wave_run_sum); a longer one is set aside and then summed by the whole
// workgroup: thread t adds elements t, t+256, t+512, ... in order into two
// interleaved partial sums (t, t+512, ... and t+256, t+768, ...), loading four
// steps ahead, and the threads' sums are combined by a fixed tree in LDS. Runs
// of different rows: no conflicts. Every thread of the workgroup calls it.
__device__ __forceinline__ void sum_long_runs(const uint32_t* __restrict__ keys_sorted,
                                              const Rgb* __restrict__ vals_sorted, uint32_t n, const uint2* lruns,
                                              uint32_t count, float* __restrict__ grad)
{
    __shared__ float part[3][kAlbBlock];
    __shared__ uint2 big[kLongPerWindow]; // runs longer than kWaveRun, ends resolved
    __shared__ uint32_t n_big;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0u) n_big = 0u;
    __syncthreads();
    for (uint32_t j = wave; j < count; j += kAlbBlock / 64u) {
        const uint2 run = lruns[j];
        const uint32_t b = run.x, row = keys_sorted[b];
        const uint32_t e = run.y != 0u ? run.y : wave_run_end(keys_sorted, n, b, row, lane);
        if (e - b > kWaveRun) {
            if (lane == 0u) big[atomicAdd(&n_big, 1u)] = make_uint2(b, e); // summed below, in slot order
            continue;
        }
        float sum[3];
        wave_run_sum(vals_sorted, b, e, lane, sum);
        if (lane == 0u) {
            grad[(size_t)row * 3 + 0] += sum[0];
            grad[(size_t)row * 3 + 1] += sum[1];
            grad[(size_t)row * 3 + 2] += sum[2];
        }
    }
    __syncthreads();
    const uint32_t nb = n_big; // each big run's sum depends on the run alone, not on its place in big[]
    for (uint32_t j = 0; j < nb; ++j) {
        const uint32_t b = big[j].x, e = big[j].y;
        const uint32_t row = keys_sorted[b];
        float s0[3] = {0.0f, 0.0f, 0.0f}, s1[3] = {0.0f, 0.0f, 0.0f};
        uint32_t i = b + t;
        for (; i + 7u * kAlbBlock < e; i += 8u * kAlbBlock) { // four steps: eight loads in flight
            Rgb v[8];
#pragma unroll
            for (uint32_t k = 0; k < 8u; ++k) v[k] = vals_sorted[i + k * kAlbBlock];
#pragma unroll
            for (uint32_t k = 0; k < 8u; k += 2u) {
                s0[0] += v[k].r; s0[1] += v[k].g; s0[2] += v[k].b;
                s1[0] += v[k + 1].r; s1[1] += v[k + 1].g; s1[2] += v[k + 1].b;
            }
        }
        for (; i + kAlbBlock < e; i += 2u * kAlbBlock) {
            const Rgb v0 = vals_sorted[i], v1 = vals_sorted[i + kAlbBlock];
            s0[0] += v0.r; s0[1] += v0.g; s0[2] += v0.b;
            s1[0] += v1.r; s1[1] += v1.g; s1[2] += v1.b;
        }
        if (i < e) {
            const Rgb v0 = vals_sorted[i];
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
            grad[(size_t)row * 3 + 0] += part[0][0];
            grad[(size_t)row * 3 + 1] += part[1][0];
            grad[(size_t)row * 3 + 2] += part[2][0];
        }
        __syncthreads(); // part[] is reused by the next run
    }
}


__global__ void __launch_bounds__(kAlbBlock) k_reduce_runs(const uint32_t* __restrict__ keys_sorted,
                                                          const Rgb* __restrict__ vals_sorted, uint32_t n,
                                                          uint32_t miss_key, float* __restrict__ grad)
{
    __shared__ uint32_t starts[kRunWindow / 32]; // bit i: window position i starts a run
    __shared__ uint2 runs[kRunWindow];           // short runs: first position, end
    __shared__ uint32_t run_row[kRunWindow];
    __shared__ uint2 lruns[kLongPerWindow];      // long runs: first position, end (0: past the window)
    const uint32_t t = threadIdx.x;
    const uint32_t w0 = blockIdx.x * kRunWindow, p0 = w0 + t * kRunItems;
    uint32_t row[kRunItems];
#pragma unroll
    for (uint32_t j = 0; j < kRunItems; ++j) row[j] = p0 + j < n ? keys_sorted[p0 + j] : 0u;
    const uint32_t before = (p0 > 0u && p0 - 1u < n) ? keys_sorted[p0 - 1u] : ~row[0];
    uint32_t mine = 0; // this thread's run starts, bit j = position p0 + j
#pragma unroll
    for (uint32_t j = 0; j < kRunItems; ++j) {
        const uint32_t prev = j == 0u ? before : row[j - 1u];
        if (p0 + j < n && prev != row[j]) mine |= 1u << j;
    }
    if (t < kRunWindow / 32u) starts[t] = 0u;
    __syncthreads();
    atomicOr(&starts[(t * kRunItems) >> 5], mine << ((t * kRunItems) & 31u)); // OR: order-free
    __syncthreads();
    uint32_t end[kRunItems];
    uint32_t n_short = 0, n_longs = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRunItems; ++j) {
        end[j] = 0u;
        if (!((mine >> j) & 1u)) continue;
        // the next run start after window position L, else past the window
        const uint32_t L = t * kRunItems + j;
        uint32_t wi = L >> 5, bits = (L & 31u) == 31u ? 0u : starts[wi] & (~0u << ((L & 31u) + 1u));
        while (bits == 0u && ++wi < kRunWindow / 32u) bits = starts[wi];
        if (row[j] >= miss_key) continue; // the misses' run
        const uint32_t p = p0 + j, wend = w0 + kRunWindow;
        uint32_t e = 0u; // 0: past the window and longer than kShortRun
        if (bits != 0u) {
            e = w0 + wi * 32u + (uint32_t)(__ffs(bits) - 1);
        } else if (wend >= n) {
            e = n;
        } else if (wend - p <= kShortRun) {
            // positions p .. wend - 1 are in the run: is it closed within kShortRun + 1 pairs?
            uint32_t differ = 0;
#pragma unroll
            for (uint32_t k = 0; k < kShortRun; ++k) {
                const uint32_t q = wend + k;
                if (q < p + kShortRun + 1u && (q >= n || keys_sorted[q < n ? q : 0u] != row[j])) differ |= 1u << k;
            }
            if (differ != 0u) e = wend + (uint32_t)(__ffs(differ) - 1);
            e = e < n ? e : n;
        }
        end[j] = e;
        if (e == 0u || e - p > kShortRun) ++n_longs;
        else ++n_short;
    }
    uint32_t total, total_long;
    uint32_t slot = block_exclusive_scan(n_short, total);
    uint32_t lslot = block_exclusive_scan(n_longs, total_long);
#pragma unroll
    for (uint32_t j = 0; j < kRunItems; ++j) {
        const uint32_t p = p0 + j;
        if (!((mine >> j) & 1u) || row[j] >= miss_key) continue;
        if (end[j] != 0u && end[j] - p <= kShortRun) {
            runs[slot] = make_uint2(p, end[j]);
            run_row[slot] = row[j];
            ++slot;
        } else {
            lruns[lslot++] = make_uint2(p, end[j]);
        }
    }
    __syncthreads();
    for (uint32_t r = t; r < total; r += kAlbBlock) {
        const uint2 run = runs[r];
        // added in order, four loads in flight at a time
        float sr = 0.0f, sg = 0.0f, sb = 0.0f;
        for (uint32_t i = run.x; i < run.y; i += 4u) {
            Rgb v[4];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k)
                if (i + k < run.y) v[k] = vals_sorted[i + k];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k)
                if (i + k < run.y) {
                    sr += v[k].r;
                    sg += v[k].g;
                    sb += v[k].b;
                }
        }
        const size_t g = (size_t)run_row[r] * 3;
        grad[g + 0] += sr;
        grad[g + 1] += sg;
        grad[g + 2] += sb;
    }
    sum_long_runs(keys_sorted, vals_sorted, n, lruns, total_long, grad);
}

// Device buffers of the deterministic reduction for up to cap_px pixels.
struct DetWorkspace {
    uint32_t *keys, *keys_sorted; // row keys: written in pixel order, sorted in place of each other
    uint32_t *pix, *pix_sorted;   // the pixel of each key
    Rgb* vals;                    // contributions, by pixel
    Rgb* vals_sorted;             // contributions in sorted order
    uint32_t* count;              // the sort's digit counts per tile (kRadix x tiles), then its kRadix digit totals
    size_t cap_px;
};

__host__ __device__ constexpr size_t sort_tiles(size_t px) { return (px + kSortTile - 1) / kSortTile; }
// words of DetWorkspace::count for px pixels
constexpr size_t count_words(size_t px) { return kRadix * sort_tiles(px) + kRadix; }
constexpr size_t run_windows(size_t px) { return (px + kRunWindow - 1) / kRunWindow; }

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

static vx_status grow(vx_albedo_state* a, size_t px)
{
    DetWorkspace& w = a->ws;
    if (px <= w.cap_px) return VX_OK;
    for (void* p : {(void*)w.keys, (void*)w.keys_sorted, (void*)w.pix, (void*)w.pix_sorted, (void*)w.vals,
                    (void*)w.vals_sorted, (void*)w.count})
        (void)hipFree(p);
    w = DetWorkspace{};
    if (hipMalloc(&w.keys, px * 4) != hipSuccess || hipMalloc(&w.keys_sorted, px * 4) != hipSuccess ||
        hipMalloc(&w.pix, px * 4) != hipSuccess || hipMalloc(&w.pix_sorted, px * 4) != hipSuccess ||
        hipMalloc(&w.vals, px * sizeof(Rgb)) != hipSuccess || hipMalloc(&w.vals_sorted, px * sizeof(Rgb)) != hipSuccess ||
        hipMalloc(&w.count, count_words(px) * 4) != hipSuccess)
        return fail(VX_ERR_OUT_OF_MEMORY, "albedo workspace allocation failed");
    w.cap_px = px;
    return VX_OK;
}

// The deterministic reduction's workspace for px pixels carved from the
// device's scratch array in 256-byte aligned slices. Fails when they do not fit.
static vx_status scratch_workspace(int device, size_t px, DetWorkspace& w)
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
    w.pix = static_cast<uint32_t*>(take(px * 4));
    w.pix_sorted = static_cast<uint32_t*>(take(px * 4));
    w.vals = static_cast<Rgb*>(take(px * sizeof(Rgb)));
    w.vals_sorted = static_cast<Rgb*>(take(px * sizeof(Rgb)));
    w.count = static_cast<uint32_t*>(take(count_words(px) * 4));
    w.cap_px = px;
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
    for (void* p : {(void*)a->base, (void*)w.keys, (void*)w.keys_sorted, (void*)w.pix, (void*)w.pix_sorted,
                    (void*)w.vals, (void*)w.vals_sorted, (void*)w.count})
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
    const uint32_t tiles = ((w + 15) / 16) * ((h + 15) / 16), per = (tiles + vx::kXcds - 1) / vx::kXcds;
    vx::k_albedo_forward<<<dim3(per * vx::kXcds), dim3(vx::kAlbBlock), lds, st>>>(v, c, w, h, per, a->base, albedo,
                                                                                  image, hit_row, hit_shade);
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
    const uint32_t tiles = ((w + 15) / 16) * ((h + 15) / 16), per = (tiles + kXcds - 1) / kXcds;
    const bool recompute = hit_row == nullptr;
    const dim3 grid(per * kXcds), block(kAlbBlock);
    if (a->voxels >= 0xFFFFFFFFull) return fail(VX_ERR_UNSUPPORTED, std::string(where) + ": too many voxels");
    const uint32_t n_voxels = (uint32_t)a->voxels;
    if (mode == VX_GRAD_ATOMIC) {
        if (recompute)
            k_albedo_adjoint<true, true><<<grid, block, lds, st>>>(v, c, w, h, per, a->base, grad_image, nullptr, nullptr,
                                                                   grad_albedo, nullptr, nullptr, nullptr, n_voxels);
        else
            k_albedo_adjoint<false, true><<<grid, block, 0, st>>>(v, c, w, h, per, a->base, grad_image, hit_row, hit_shade,
                                                                  grad_albedo, nullptr, nullptr, nullptr, n_voxels);
        if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": launch failed");
        return VX_OK;
    }
    // deterministic: sort (row, pixel) by row, then sum each run in a fixed order
    if (px > 0x7FFFFFFFu) return fail(VX_ERR_INVALID_ARG, std::string(where) + ": deterministic mode needs w*h < 2^31");
    // misses get key V (the voxel count) and sort last; the sort looks at the key bits V needs
    const uint32_t miss_key = n_voxels;
    uint32_t end_bit = 1;
    while (end_bit < 32 && (miss_key >> end_bit) != 0u) ++end_bit;
    DetWorkspace ws{};
    if (noalloc) {
        if (vx_status g = scratch_workspace(s->device, px, ws); g != VX_OK) return g;
    } else {
        if (vx_status g = grow(a, px); g != VX_OK) return g;
        ws = a->ws;
    }
    if (recompute)
        k_albedo_adjoint<true, false><<<grid, block, lds, st>>>(v, c, w, h, per, a->base, grad_image, nullptr, nullptr,
                                                                nullptr, ws.keys, ws.vals, ws.pix, miss_key);
    else
        k_albedo_adjoint<false, false><<<grid, block, 0, st>>>(v, c, w, h, per, a->base, grad_image, hit_row, hit_shade,
                                                               nullptr, ws.keys, ws.vals, ws.pix, miss_key);
    if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": launch failed");
    // 8 key bits per pass, ping-ponging between the two key and pixel buffers
    const uint32_t n = (uint32_t)px, n_tiles = (uint32_t)sort_tiles(px);
    uint32_t *kin = ws.keys, *vin = ws.pix, *kout = ws.keys_sorted, *vout = ws.pix_sorted;
    uint32_t* total = ws.count + (size_t)kRadix * n_tiles;
    for (uint32_t shift = 0; shift < end_bit; shift += 8) {
        k_sort_count<<<dim3(n_tiles), dim3(kSortThreads), 0, st>>>(kin, n, shift, n_tiles, ws.count);
        k_sort_scan<<<dim3(kRadix), dim3(kSortThreads), 0, st>>>(ws.count, n_tiles, total);
        if (shift + 8u >= end_bit) // the last pass also puts the contributions in sorted order
            k_sort_scatter<true><<<dim3(n_tiles), dim3(kSortThreads), 0, st>>>(kin, vin, n, shift, n_tiles, ws.count,
                                                                               total, kout, vout, ws.vals,
                                                                               ws.vals_sorted);
        else
            k_sort_scatter<false><<<dim3(n_tiles), dim3(kSortThreads), 0, st>>>(kin, vin, n, shift, n_tiles, ws.count,
                                                                                total, kout, vout, nullptr, nullptr);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (hipGetLastError() != hipSuccess) return fail(VX_ERR_HIP, std::string(where) + ": sort failed");
    // the sorted keys are in kin, their contributions in vals_sorted; one workgroup per window of positions
    const uint32_t n_windows = (uint32_t)run_windows(px);
    k_reduce_runs<<<dim3(n_windows), block, 0, st>>>(kin, ws.vals_sorted, n, miss_key, grad_albedo);
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
