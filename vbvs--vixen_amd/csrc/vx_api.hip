// vx_api.hip -- the C ABI of libvixen_hip.so: contexts, scene upload and the
// trace/render launches for gfx950.
//
// Launch geometry (plain kernels): 256-lane workgroups (4 x wave64), one ray per
// lane. Workgroup ids are remapped so that the 8 XCDs each own a contiguous
// band of tiles/ray chunks (the hardware deals workgroup i to XCD i % 8), which
// keeps neighbouring rays -- and the descriptors they share -- in one XCD's L2.
// A prefix of the descriptor array (the top breadth-first levels) can be staged
// in LDS per workgroup; vx_scene_set_lds_depth picks how many levels. Depth 0
// (the default) launches kernels compiled without staging: their descriptor
// fetch is a plain global load, which measures faster on gfx950 than the flat
// load a runtime LDS-or-global choice compiles to (the top levels stay in L2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <new>
#include <string>
#include <vector>

#include "vx_internal.hpp"

namespace vx {

static thread_local std::string g_last_error;

vx_status fail(vx_status code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}
void clear_error() { g_last_error.clear(); }

constexpr uint32_t kBlock = 256;

struct CamParams {
    float pos[3], dir[3], up[3], right[3];
    float tan_half, aspect;
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Occupancy bit mask: word w of a brick holds voxels 32w..32w+31 (bit j = voxel 32w+j).
__global__ void __launch_bounds__(kBlock) k_build_occ(const uint32_t* __restrict__ brick_data, uint64_t words,
                                                      uint32_t* __restrict__ occ)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= words) return;
    const uint4* src = reinterpret_cast<const uint4*>(brick_data + i * 32u);
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint4 v = src[q];
        m |= (v.x != 0u ? 1u : 0u) << (4 * q + 0);
        m |= (v.y != 0u ? 1u : 0u) << (4 * q + 1);
        m |= (v.z != 0u ? 1u : 0u) << (4 * q + 2);
        m |= (v.w != 0u ? 1u : 0u) << (4 * q + 3);
    }
    occ[i] = m;
}

// Per brick, the box of its occupied voxels from the 16 occupancy words (bit b
// of word w is voxel 32w + b = x + 8y + 64z): lo x|y|z in bits 0-8, hi x|y|z in
// bits 9-17. An empty brick gets the whole brick (the pretest then rejects
// nothing and the DDA finds nothing).
__global__ void __launch_bounds__(kBlock) k_build_box(const uint32_t* __restrict__ occ, uint64_t nb,
                                                      uint32_t* __restrict__ box)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nb) return;
    const uint4* m = reinterpret_cast<const uint4*>(occ + i * 16u);
    uint32_t xs = 0u, ys = 0u, zs = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = m[q];
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int w = 4 * q + j; // z = w / 2, rows y = (w & 1) * 4 + byte
            const uint32_t word = w4[j];
            zs |= (word != 0u ? 1u : 0u) << (w >> 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t row = (word >> (8 * k)) & 0xFFu;
                xs |= row;
                ys |= (row != 0u ? 1u : 0u) << ((w & 1) * 4 + k);
            }
        }
    }
    if (zs == 0u) {
        box[i] = 7u << 9 | 7u << 12 | 7u << 15;
        return;
    }
    const uint32_t lx = __builtin_ctz(xs), ly = __builtin_ctz(ys), lz = __builtin_ctz(zs);
    const uint32_t hx = 31u - __builtin_clz(xs), hy = 31u - __builtin_clz(ys), hz = 31u - __builtin_clz(zs);
    box[i] = lx | ly << 3 | lz << 6 | hx << 9 | hy << 12 | hz << 15;
}

// Stripe lengths: k_trace's chunks are 256 consecutive rays, k_render's 16x16 pixel tiles.
constexpr uint32_t kTraceStripe = 4;

// Every lane of the workgroup copies part of the descriptor prefix into LDS.
template <bool STAGED>
__device__ __forceinline__ void stage_prefix(const SceneView& s)
{
    if (!STAGED) return;
    for (uint32_t i = threadIdx.x; i < s.lds_count; i += blockDim.x) vx_lds_nodes[i] = s.nodes[i];
    __syncthreads();
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v)
{
    unsigned long long x = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63u) == 0u) atomicAdd(dst, x);
}

// Rays of one 64-ray block (rays 64b .. 64b+63 below n, one per lane) that all
// start at one origin (bit-equal to the block's first ray's): a pinhole
// camera's rays. Such a block is traced one ray per lane in lockstep, leaves
// visited at once; any other block postpones leaf visits (trace_ray's POSTPONE)
// or, when SPLIT, is left to k_trace_refill.
__device__ __forceinline__ bool block_shares_origin(const float o[3])
{
    const bool same = f2u(o[0]) == __builtin_amdgcn_readfirstlane(f2u(o[0])) &&
                      f2u(o[1]) == __builtin_amdgcn_readfirstlane(f2u(o[1])) &&
                      f2u(o[2]) == __builtin_amdgcn_readfirstlane(f2u(o[2]));
    return __ballot(!same) == 0ull;
}

// Flags word that k_trace<..., SPLIT> writes into the hit record of the first
// ray of each part of a block it leaves to k_trace_refill: face-axis field 3,
// which no record of a traced ray holds (axis 0-2). k_trace_refill deals each
// wave a range of positions in a scattered block order (below), so a block is
// one part, or two when a range boundary falls inside it. A part's first ray is
// traced only by the wave whose range holds the part, after that wave has read
// the ray's record; every such record is written by the first launch, with this
// mark or with the ray's own record, so the refill launch classifies a part by
// one word instead of re-reading its origins.
constexpr uint32_t kLeftToRefill = 3u << 1;

// k_trace_refill's dealing, known to the first launch so it can mark the parts:
// the range of positions per wave and the inverse of the block order's step
// modulo the block count (block b is dealt at positions 64 k .. 64 k + 63,
// k = b * inv_step mod n_blocks).
struct RefillDeal {
    uint64_t range, inv_step;
};

template <bool COUNT, bool STAGED, bool LOD = false, bool SPLIT = false>
__global__ void __launch_bounds__(kBlock) k_trace(SceneView s, const vx_ray* __restrict__ rays, uint64_t n,
                                                  uint32_t per_xcd, vx_hit* __restrict__ hits,
                                                  unsigned long long* __restrict__ counters, RefillDeal deal)
{
    stage_prefix<STAGED>(s);
    const uint64_t i = (uint64_t)xcd_stripe_chunk<kTraceStripe>(per_xcd) * kBlock + threadIdx.x;
    TraceCounts cnt{0u, 0u, 0u, 0u};
    if (i < n) {
        // a wave holds one aligned 64-ray block; its origins decide how it is traced
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 a = rp[0];
        const float o[3] = {a.x, a.y, a.z};
        const float tmin = a.w;
        const bool shared = block_shares_origin(o);
        if (SPLIT && !shared) { // wave-uniform: k_trace_refill traces this block
            const uint64_t k = ((i >> 6) * deal.inv_step) % ((n + 63u) >> 6); // below 2^32 each
            const uint64_t p0 = k << 6, q = (p0 + deal.range - 1u) / deal.range * deal.range;
            const uint64_t lane = i & 63u;
            if (lane == 0u || (q < p0 + 64u && lane == q - p0)) hits[i].flags = kLeftToRefill;
            return;
        }
        const float4 b = rp[1];
        const float d[3] = {b.x, b.y, b.z};
        const float tmax = b.w;
        vx_hit h;
        if (shared) trace_ray<COUNT, STAGED, false, LOD>(s, o, d, tmin, tmax, h, cnt);
        else trace_ray<COUNT, STAGED, true, LOD>(s, o, d, tmin, tmax, h, cnt);
        store_hit(hits, i, h);
    }
    if (COUNT) {
        wave_add_u64(&counters[0], i < n ? 1u : 0u);
        wave_add_u64(&counters[1], cnt.iters);
        wave_add_u64(&counters[2], cnt.leaves);
        wave_add_u64(&counters[3], cnt.voxels);
        wave_add_u64(&counters[4], cnt.hits);
    }
}

// vx_trace's second launch: the 64-ray blocks whose rays do not share one
// origin (random, bounce or other incoherent rays), traced with lane refill.
// One ray per lane for a whole block wastes most of the wave on incoherent
// rays: rays of a block end after very different numbers of iterations, so
// few lanes stay active. Here a resident grid of waves keeps 64 rays in
// flight per wave: whenever at least refill_min lanes are idle (between ESVO
// iterations) they take the next rays of the wave's range (ranked by mbcnt).
// The rays are dealt as one contiguous range of `range` rays per wave, equal
// in length, in a scattered block order: position p of the dealt order is ray
// p % 64 of block (p / 64) * step mod n_blocks (step coprime to n_blocks, so a
// bijection). Clusters of incoherent blocks -- a set of camera rays followed
// by a set of random rays, say -- then spread evenly over the waves, with no
// atomics or device state (the launch stays capturable and re-entrant), while
// each wave still takes its rays 64 neighbours at a time. A block may be
// split between two waves' ranges; each wave reads the hit record of the first
// ray of its part before it traces any ray of that part (no other wave traces
// that ray) and skips the part unless k_trace<..., SPLIT> marked it
// kLeftToRefill. Leaf visits wait
// until 3/4 of the live lanes wait at one (trace_ray's POSTPONE). Each lane
// executes exactly the reference's iteration sequence, so records equal
// k_trace's bit for bit.
template <bool STAGED, bool LOD>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7, 7)))
    k_trace_refill(SceneView s, const vx_ray* __restrict__ rays, uint64_t n, uint64_t range, uint64_t step,
                   uint32_t refill_min, vx_hit* __restrict__ hits)
{
    stage_prefix<STAGED>(s);
    const uint64_t v = (uint64_t)blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6);
    const uint64_t n_blocks = (n + 63u) >> 6;
    const uint64_t slots = n_blocks << 6; // dealt positions: every block counts 64, its rays past n are empty
    uint64_t pos = min(v * range, slots);        // next dealt position not yet classified
    const uint64_t last = min(pos + range, slots);
    uint64_t cur = 0, end = 0;                   // rays of the current block not yet handed out
    bool drained = false;

    enum : int { IDLE = 0, ESVO = 1, LEAF = 2 };
    int mode = IDLE;
    uint64_t ray_id = 0;
    RayState r;
    vx_hit h;
    TraceCounts cnt{0u, 0u, 0u, 0u};
    for (;;) {
        uint64_t idle = __ballot(mode == IDLE);
        if (!drained && (uint32_t)__popcll(idle) >= refill_min) {
            while (idle != 0ull) {
                while (cur == end) { // the next block of the range that k_trace leaves to this launch
                    if (pos >= last) {
                        drained = true;
                        break;
                    }
                    const uint64_t stop = min((pos | 63u) + 1u, last); // the dealt block's part in the range
                    const uint64_t b = ((pos >> 6) * step) % n_blocks;
                    const uint64_t lo = (b << 6) + (pos & 63u), hi = min((b << 6) + ((stop - 1u) & 63u) + 1u, n);
                    pos = stop;
                    if (lo >= hi || hits[lo].flags != kLeftToRefill) continue; // wave-uniform
                    cur = lo;
                    end = hi;
                }
                if (drained) break;
                const uint64_t take = min((uint64_t)__popcll(idle), end - cur);
                const uint32_t rank =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                if (mode == IDLE && rank < take) {
                    ray_id = cur + rank;
                    float o[3], d[3], tmin, tmax;
                    load_ray(rays, ray_id, o, d, tmin, tmax);
                    if (init_ray(s, o, d, tmin, tmax, r)) {
                        rec_load(s, h);
                        store_hit(hits, ray_id, h);
                    } else {
                        mode = ESVO;
                    }
                }
                cur += take;
                idle = __ballot(mode == IDLE);
            }
        }
        const uint64_t live = __ballot(mode != IDLE);
        if (live == 0ull) {
            if (drained) break;
            continue;
        }
        bool done = false;
        if (mode == ESVO) {
            LeafRef lf;
            const int st = esvo_step<false, STAGED, false, LOD>(s, r, cnt, lf);
            if (st == STEP_DONE) {
                done = true;
            } else if (st == STEP_LEAF) {
                *lane_word(s, s.aux_base, 11) = lf.brick;
                *lane_word(s, s.aux_base, 12) = lf.lidx;
                mode = LEAF;
            }
        }
        const uint32_t waiting = (uint32_t)__popcll(__ballot(mode == LEAF));
        if (waiting * 4u >= (uint32_t)__popcll(live) * 3u && mode == LEAF) {
            int bc[3];
            brick_cell(s, r, bc);
            if (visit_leaf<false>(s, r, *lane_word(s, s.aux_base, 11), bc, *lane_word(s, s.aux_base, 12), cnt))
                done = true;
            else if (resume_after_leaf(s, r))
                done = true;
            else
                mode = ESVO;
        }
        if (done) {
            rec_load(s, h);
            store_hit(hits, ray_id, h);
            mode = IDLE;
        }
    }
}

// Pinhole ray of pixel (px, py): getRayDir (shaders/RayGeneration.glsl:61-71) at
// the pixel centre (VoxelRayMarch_Compressed.comp:501-503). Unnormalised; the
// trace normalises it.
__device__ __forceinline__ void camera_dir(const CamParams& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h,
                                           float d[3], float& v)
{
    const float u = ((float)px + 0.5f) / (float)w;
    v = ((float)py + 0.5f) / (float)h;
    const float ndcx = u * 2.0f - 1.0f;
    const float ndcy = -(v * 2.0f - 1.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a = ((c.right[i] * ndcx) * c.tan_half) * c.aspect;
        const float b = (c.up[i] * ndcy) * c.tan_half;
        d[i] = (c.dir[i] + a) + b;
    }
}

__device__ __forceinline__ uint32_t to_u8(float x)
{
    x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    return (uint32_t)(x * 255.0f + 0.5f);
}

// computeLighting's fixed light, normalize(vec3(1, 1, -1)), folded to the
// correctly rounded 1/sqrt(3) as a compiler folds the constant.
constexpr float kLight = 0.57735026918962576f;

// decodeDXT1Color as the compressed shader evaluates it (shaders/Compression.glsl:28-40,60-75):
// c0 from {2^-32, 0, 2*2^-32/3, 2^-32/3}, c1 = 2^-32 - c0, channel = c0 * float(head << k) + c1 * float(head << m).
__device__ __forceinline__ void bc1_decode_shader(uint2 block, uint32_t texel, float rgb[3])
{
    const float S = 0x1p-32f;
    const uint32_t sel = (block.y >> (2u * texel)) & 3u;
    const float c0 = sel == 0u ? S : (sel == 1u ? 0.0f : (sel == 2u ? (2.0f * S) / 3.0f : (1.0f * S) / 3.0f));
    const float c1 = S - c0;
    const uint32_t h = block.x;
    rgb[0] = c0 * (float)(h << 16) + c1 * (float)h;
    rgb[1] = c0 * (float)(h << 21) + c1 * (float)(h << 5);
    rgb[2] = c0 * (float)(h << 27) + c1 * (float)(h << 11);
}

// decodeDXTNormal (shaders/Compression.glsl:115-177): raw base normal plus the
// U and V axes weighted by {-1, -1/3, 1/3, 1}. Unnormalised.
__device__ __forceinline__ void dxt_normal_decode(uint4 blk, uint32_t texel, float n[3])
{
    const uint32_t hb = blk.x, bits_u = blk.y, huv = blk.z, bits_v = blk.w;
    const int32_t sgn = (int32_t)hb >> 31;
    const uint32_t axis = (hb >> 29) & 3u;
    const float t = (float)(sgn ^ 0x7FFFFFFF), u = (float)(int32_t)(hb << 3), v = (float)(int32_t)(hb << 18);
    const float bx = axis == 0u ? t : (axis == 1u ? v : u);
    const float by = axis == 0u ? u : (axis == 1u ? t : v);
    const float bz = axis == 0u ? v : (axis == 1u ? u : t);
    float ax[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int32_t sh = (int32_t)(huv << (16 - 16 * k));
        const int e = (sh >> 16) & 15;
        const float scale = __uint_as_float((uint32_t)(e + (3 - 13) + 127) << 23); // exp2(e - 10), e in [0,15]
        ax[k][0] = (float)sh * scale;
        ax[k][1] = (float)(int32_t)((uint32_t)sh << 4) * scale;
        ax[k][2] = (float)(int32_t)((uint32_t)sh << 8) * scale;
    }
    const uint32_t su = (bits_u >> (2u * texel)) & 3u, sv = (bits_v >> (2u * texel)) & 3u;
    const float c13 = 1.0f / 3.0f;
    const float cu = su == 0u ? -1.0f : (su == 1u ? -c13 : (su == 2u ? c13 : 1.0f));
    const float cv = sv == 0u ? -1.0f : (sv == 1u ? -c13 : (sv == 2u ? c13 : 1.0f));
    n[0] = (bx + ax[0][0] * cu) + ax[1][0] * cv;
    n[1] = (by + ax[0][1] * cu) + ax[1][1] * cv;
    n[2] = (bz + ax[0][2] * cu) + ax[1][2] * cv;
}

__device__ __forceinline__ void normalize3(float v[3])
{
    const float inv = 1.0f / __builtin_sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
}

// The compressed shader's shading of a hit (VoxelRayMarch_Compressed.comp:154-165,
// 239-256,541-542; shaders/Lighting.glsl:19-28): colour from the shader's BC1
// decode, normal from the DXT normal block blended with the DDA step normal, or
// the ray's dominant-axis normal when the DDA took no step in the hit brick.
// du is the normalised world direction.
__device__ __forceinline__ void shade_compressed(const SceneView& s, const vx_hit& h, const float du[3],
                                                 float col[3])
{
    float rgb[3];
    bc1_decode_shader(s.bc1[(size_t)h.brick * 32u + (h.voxel >> 4)], h.voxel & 15u, rgb);
    const uint32_t last = (h.flags >> 4) & 3u;
    float n[3] = {0.0f, 0.0f, 0.0f};
    if (last == 0u) {
        const float a0 = __builtin_fabsf(du[0]), a1 = __builtin_fabsf(du[1]), a2 = __builtin_fabsf(du[2]);
        const int ax = (a0 > a1 && a0 > a2) ? 0 : (a1 > a2 ? 1 : 2);
        const float d = ax == 0 ? du[0] : (ax == 1 ? du[1] : du[2]);
        const float sg = d > 0.0f ? -1.0f : (d < 0.0f ? 1.0f : 0.0f); // -sign(d)
        n[0] = ax == 0 ? sg : 0.0f;
        n[1] = ax == 1 ? sg : 0.0f;
        n[2] = ax == 2 ? sg : 0.0f;
    } else {
        dxt_normal_decode(s.nrm[(size_t)h.brick * 32u + (h.voxel >> 4)], h.voxel & 15u, n);
        normalize3(n);
        const float d = last == 1u ? du[0] : (last == 2u ? du[1] : du[2]);
        const float dda = d < 0.0f ? 1.0f : -1.0f; // -step, step = sign(d) with 0 -> 1
        n[0] = n[0] + (last == 1u ? dda : 0.0f) * 0.5f;
        n[1] = n[1] + (last == 2u ? dda : 0.0f) * 0.5f;
        n[2] = n[2] + (last == 3u ? dda : 0.0f) * 0.5f;
        normalize3(n);
    }
    const float dt = (n[0] * kLight + n[1] * kLight) + n[2] * -kLight;
    const float k = 0.3f + (dt < 0.0f ? 0.0f : dt) * 0.7f;
    col[0] = rgb[0] * k;
    col[1] = rgb[1] * k;
    col[2] = rgb[2] * k;
}

// Shading of vx_render's RGBA8 image, chosen per launch so the hits-only
// kernel carries none of it.
enum : int { SHADE_NONE = 0, SHADE_FACE = 1, SHADE_COMPRESSED = 2 };

// 16x16 pixel tile per workgroup; each wave covers an 8x8 sub-tile.
template <bool STAGED, int SHADE>
__device__ __forceinline__ void render_tile(const SceneView& s, const CamParams& c, uint32_t w, uint32_t h,
                                            float tmax, uint32_t tiles_x, uint32_t ntiles, uint32_t per_xcd,
                                            vx_hit* __restrict__ hits, uint32_t* __restrict__ rgba)
{
    stage_prefix<STAGED>(s);
    const uint32_t tile = xcd_stripe_chunk<kRenderStripe>(per_xcd);
    if (tile >= ntiles) return;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t px = (tile % tiles_x) * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * 16u + (wave >> 1) * 8u + (lane >> 3);
    if (px >= w || py >= h) return;
    float d[3], v;
    camera_dir(c, px, py, w, h, d, v);
    vx_hit hr;
    TraceCounts cnt{0u, 0u, 0u, 0u};
    const bool hit = trace_ray<false, STAGED, false>(s, c.pos, d, 0.0f, tmax, hr, cnt);
    const uint64_t i = (uint64_t)py * w + px;
    if (hits) store_hit(hits, i, hr);
    if (SHADE != SHADE_NONE) {
        float col[3];
        if (SHADE == SHADE_COMPRESSED && hit) {
            float du[3];
            const float inv = 1.0f / __builtin_sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
            for (int i = 0; i < 3; ++i) du[i] = d[i] * inv; // as setup_ray normalises it
            shade_compressed(s, hr, du, col);
        } else if (hit) {
            // computeLighting (shaders/Lighting.glsl:19-28) with the entry-face normal
            const uint32_t axis = (hr.flags >> 1) & 3u;
            const float sgn = (hr.flags & 8u) ? 1.0f : -1.0f;
            const float l = 0.57735026918962576f; // normalize(1,1,-1) components have magnitude 1/sqrt(3)
            const float ldot = axis == 2u ? -sgn * l : sgn * l;
            const float k = 0.3f + (ldot > 0.0f ? ldot : 0.0f) * 0.7f;
            col[0] = hr.rgb[0] * k;
            col[1] = hr.rgb[1] * k;
            col[2] = hr.rgb[2] * k;
        } else {
            // sky (VoxelRayMarch_Compressed.comp:544)
            const float g = 1.0f - v * 0.5f;
            col[0] = 0.5f * g;
            col[1] = 0.7f * g;
            col[2] = 1.0f * g;
        }
        rgba[i] = to_u8(col[0]) | (to_u8(col[1]) << 8) | (to_u8(col[2]) << 16) | 0xFF000000u;
    }
}

// Hits-only and entry-face shading fit 64 VGPRs: asked for 8 waves per SIMD
// (measured 3% faster than 7 on the 1080p cave; the LDS columns of 8
// workgroups still fit). The compressed shading keeps its natural register
// budget (at 8 waves it would spill).
template <bool STAGED, int SHADE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_render(SceneView s, CamParams c, uint32_t w, uint32_t h, float tmax, uint32_t tiles_x, uint32_t ntiles,
         uint32_t per_xcd, vx_hit* __restrict__ hits, uint32_t* __restrict__ rgba)
{
    static_assert(SHADE != SHADE_COMPRESSED, "compressed shading runs in k_render_compressed");
    render_tile<STAGED, SHADE>(s, c, w, h, tmax, tiles_x, ntiles, per_xcd, hits, rgba);
}

template <bool STAGED>
__global__ void __launch_bounds__(kBlock)
k_render_compressed(SceneView s, CamParams c, uint32_t w, uint32_t h, float tmax, uint32_t tiles_x, uint32_t ntiles,
                    uint32_t per_xcd, vx_hit* __restrict__ hits, uint32_t* __restrict__ rgba)
{
    render_tile<STAGED, SHADE_COMPRESSED>(s, c, w, h, tmax, tiles_x, ntiles, per_xcd, hits, rgba);
}

__global__ void __launch_bounds__(kBlock) k_camera_rays(CamParams c, uint32_t w, uint32_t h, float tmax,
                                                        vx_ray* __restrict__ rays)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (uint64_t)w * h) return;
    const uint32_t px = (uint32_t)(i % w), py = (uint32_t)(i / w);
    float d[3], v;
    camera_dir(c, px, py, w, h, d, v);
    float4* dst = reinterpret_cast<float4*>(rays + i);
    dst[0] = make_float4(c.pos[0], c.pos[1], c.pos[2], 0.0f);
    dst[1] = make_float4(d[0], d[1], d[2], tmax);
}

// One cosine-weighted diffuse bounce per primary hit: origin on the entry face
// of the hit voxel, pushed 1e-3 voxel off the surface along the face normal,
// direction sqrt(u1)-distributed about the normal (Malley's method) from a
// counter-based hash of (seed, ray index). Misses get a zero direction, which
// vx_trace reports as a miss.
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void __launch_bounds__(kBlock) k_bounce_rays(const vx_ray* __restrict__ prim, const vx_hit* __restrict__ hits,
                                                        uint64_t n, uint32_t seed, float k,
                                                        vx_ray* __restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float o[3], d[3], tmin, tmax;
    load_ray(prim, i, o, d, tmin, tmax);
    const vx_hit h = hits[i];
    float4* dst = reinterpret_cast<float4*>(out + i);
    // a miss, or a record whose face axis is not 0-2 (not one the trace writes),
    // gives a zero-direction ray, which the trace treats as invalid
    if (!(h.flags & 1u) || ((h.flags >> 1) & 3u) == 3u) {
        dst[0] = make_float4(o[0], o[1], o[2], 0.0f);
        dst[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float inv = 1.0f / __builtin_sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const uint32_t axis = (h.flags >> 1) & 3u;
    const float sgn = (h.flags & 8u) ? 1.0f : -1.0f;
    float nrm[3] = {0.0f, 0.0f, 0.0f};
    nrm[axis] = sgn;
    const float eps = 1e-3f / k; // 1e-3 voxel in world units
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = o[a] + d[a] * inv * h.t + nrm[a] * eps;
    const uint32_t r1 = mix32((uint32_t)i * 2u + 1u + mix32(seed ^ (uint32_t)(i >> 32)));
    const uint32_t r2 = mix32(r1 ^ 0x9e3779b9u);
    const float u1 = (float)(r1 >> 8) * 0x1p-24f, u2 = (float)(r2 >> 8) * 0x1p-24f;
    // cosine-weighted hemisphere (Malley): a uniform point on the unit disk
    // (radius sqrt(u1), azimuth 2 pi u2) lifted onto the hemisphere about the normal
    const float r = __builtin_sqrtf(u1), phi = 6.28318530717958648f * u2;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const float a1 = r * cp, a2 = r * sp, up = __builtin_sqrtf(1.0f - u1);
    const uint32_t t1 = (axis + 1u) % 3u, t2 = (axis + 2u) % 3u;
    float dir[3];
    dir[axis] = sgn * up;
    dir[t1] = a1;
    dir[t2] = a2;
    dst[0] = make_float4(p[0], p[1], p[2], 0.0f);
    dst[1] = make_float4(dir[0], dir[1], dir[2], 3.0e38f);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static CamParams make_cam(const vx_camera* cam)
{
    CamParams c;
    for (int i = 0; i < 3; ++i) {
        c.pos[i] = cam->pos[i];
        c.dir[i] = cam->dir[i];
        c.up[i] = cam->up[i];
        c.right[i] = cam->right[i];
    }
    // tan(radians(fov * 0.5)), radians(x) = x * pi/180 (RayGeneration.glsl:62)
    c.tan_half = std::tan((cam->fov_deg * 0.5f) * 0.017453292519943295f);
    c.aspect = cam->aspect;
    return c;
}

static int ilog2(uint32_t v)
{
    int l = 0;
    while ((1u << l) < v) ++l;
    return l;
}

// Walks the tree from the root and checks every pointer the traversal will
// follow: internal levels carry no leaf bits, the brick-parent level carries
// leafMask == validMask, child blocks and brick ids are in range. Also counts
// the descriptors of the top four levels (for LDS staging).
static vx_status validate_tree(const vx_scene_desc* d, int levels, uint32_t level_prefix[4])
{
    const uint32_t* n = d->esvo_nodes;
    const uint64_t nc = d->node_count;
    struct Item {
        uint64_t idx;
        int depth;
    };
    uint64_t per_depth[4] = {1, 0, 0, 0};
    std::vector<Item> todo;
    todo.push_back({0, 0});
    uint64_t visits = 0;
    const uint64_t max_visits = nc * 8 + 16;
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        if (++visits > max_visits) return fail(VX_ERR_INVALID_SCENE, "tree has more paths than descriptors allow");
        const uint32_t w0 = n[2 * it.idx], w1 = n[2 * it.idx + 1];
        const uint32_t valid = (w0 >> 16) & 0xFFu, leaf = (w0 >> 24) & 0xFFu;
        if (valid == 0) continue;
        const uint64_t cp = (w0 & 0x8000u) ? (w1 & 0xFFFFFFu) : (w0 & 0x7FFFu);
        const bool last = it.depth == levels - 1;
        if (last ? (leaf != valid) : (leaf != 0))
            return fail(VX_ERR_INVALID_SCENE, "leaf bits must mark exactly the children of brick-parent nodes (node " +
                                                  std::to_string(it.idx) + ")");
        const uint32_t cnt = static_cast<uint32_t>(__builtin_popcount(valid));
        if (cp + cnt > nc)
            return fail(VX_ERR_INVALID_SCENE, "child block of node " + std::to_string(it.idx) + " out of range");
        if (it.depth + 1 < 4) per_depth[it.depth + 1] += cnt;
        for (uint32_t c = 0; c < cnt; ++c) {
            if (last) {
                const uint32_t brick = n[2 * (cp + c) + 1] & 0xFFFFFFu;
                if (brick != 0xFFFFFFu && brick >= d->brick_count)
                    return fail(VX_ERR_INVALID_SCENE, "brick id out of range at descriptor " + std::to_string(cp + c));
            } else {
                todo.push_back({cp + c, it.depth + 1});
            }
        }
    }
    uint64_t acc = 0;
    for (int k = 0; k < 4; ++k) {
        acc += per_depth[k];
        level_prefix[k] = static_cast<uint32_t>(acc < nc ? acc : nc);
    }
    return VX_OK;
}

static vx_status scene_validate_impl(const vx_scene_desc* d, uint32_t level_prefix[4])
{
    if (!d || !d->config || !d->esvo_nodes) return fail(VX_ERR_INVALID_ARG, "vx_scene_validate: null argument");
    const vx_octree_config& c = *d->config;
    const uint32_t B = static_cast<uint32_t>(c.bricks_per_axis);
    if (c.bricks_per_axis < 2 || c.bricks_per_axis > 512 || (B & (B - 1)) != 0)
        return fail(VX_ERR_INVALID_SCENE, "bricks_per_axis must be a power of two in [2, 512]");
    const int levels = ilog2(B);
    if (c.brick_size != 8 || c.brick_depth_levels != 3 || c.user_max_levels != levels + 3 || c.esvo_max_scale != 22)
        return fail(VX_ERR_INVALID_SCENE, "config must describe 8^3 bricks under log2(bricks_per_axis) levels, esvo scale 22");
    if (levels - 1 > VX_STACK) return fail(VX_ERR_UNSUPPORTED, "tree deeper than the traversal stack");
    const float sx = c.grid_max[0] - c.grid_min[0], sy = c.grid_max[1] - c.grid_min[1], sz = c.grid_max[2] - c.grid_min[2];
    if (!(sx > 0.0f) || sx != sy || sx != sz || !std::isfinite(sx))
        return fail(VX_ERR_INVALID_SCENE, "grid bounds must be a non-empty cube");
    if (d->node_count < 1 || d->node_count > 0xFFFFFFFFull) return fail(VX_ERR_INVALID_SCENE, "node_count out of range");
    if (d->node_count >= (1ull << 29))
        return fail(VX_ERR_UNSUPPORTED, "descriptor array of 4 GiB or more (kernels address it with 32-bit offsets)");
    if (d->brick_count >= 0xFFFFFFull) return fail(VX_ERR_INVALID_SCENE, "brick_count exceeds 24-bit ids");
    if (d->brick_count > 0 && !d->brick_data) return fail(VX_ERR_INVALID_ARG, "brick_data is null");
    return validate_tree(d, levels, level_prefix);
}

// device: the buffers in d are device pointers the scene borrows (the config
// stays a host pointer); otherwise host pointers the scene copies.
static vx_status scene_create(vx_ctx* ctx, const vx_scene_desc* d, vx_scene** out, bool device)
{
    if (!ctx || !d || !out) return fail(VX_ERR_INVALID_ARG, "vx_scene_create: null argument");
    *out = nullptr;
    uint32_t level_prefix[4];
    if (!device) {
        if (vx_status st = scene_validate_impl(d, level_prefix); st != VX_OK) return st;
    } else {
        if (!d->esvo_nodes || !d->config) return fail(VX_ERR_INVALID_ARG, "vx_scene_create_device: null argument");
        auto misaligned = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) != 0; };
        if (misaligned(d->esvo_nodes, 8) || misaligned(d->brick_data, 16) || misaligned(d->bc1_colors, 8) ||
            misaligned(d->dxt_normals, 16))
            return fail(VX_ERR_INVALID_ARG, "vx_scene_create_device: descriptors, BC1 blocks need 8-byte and "
                                            "brick data, normal blocks 16-byte alignment");
        if (d->node_count < 1 || d->node_count >= (1ull << 29))
            return fail(VX_ERR_INVALID_SCENE, "node_count out of range");
        // the tree checks read the descriptors: one copy of them to the host
        std::vector<uint32_t> host(d->node_count * 2);
        VX_DEVICE_GUARD(ctx->device);
        if (hipMemcpy(host.data(), d->esvo_nodes, d->node_count * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(VX_ERR_INVALID_ARG, "vx_scene_create_device: esvo_nodes is not readable device memory");
        vx_scene_desc hd = *d;
        hd.esvo_nodes = host.data();
        if (vx_status st = scene_validate_impl(&hd, level_prefix); st != VX_OK) return st;
    }
    const vx_octree_config& c = *d->config;
    const uint32_t B = static_cast<uint32_t>(c.bricks_per_axis);
    const float sx = c.grid_max[0] - c.grid_min[0];

    VX_DEVICE_GUARD(ctx->device);

    auto* s = new (std::nothrow) vx_scene{};
    if (!s) return fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_create: out of memory");
    s->device = ctx->device;
    s->cu_count = ctx->cu_count;
    s->node_count = d->node_count;
    s->brick_count = d->brick_count;
    s->cfg = c;
    s->borrowed = device;
    std::memcpy(s->level_prefix, level_prefix, sizeof level_prefix);
    auto cleanup = [&](vx_status st) {
        if (!s->borrowed) {
            (void)hipFree(s->nodes);
            (void)hipFree(s->brick_data);
            (void)hipFree(s->bc1);
            (void)hipFree(s->nrm);
        }
        (void)hipFree(s->occ);
        (void)hipFree(s->box);
        (void)hipFree(s->palette);
        delete s;
        return st;
    };
    const uint64_t nb = d->brick_count;
    const size_t node_bytes = d->node_count * 8, brick_bytes = std::max<uint64_t>(1, nb) * 2048,
                 occ_bytes = std::max<uint64_t>(1, nb) * 64, bc1_bytes = std::max<uint64_t>(1, nb) * 256,
                 nrm_bytes = std::max<uint64_t>(1, nb) * 512;
    if (hipMalloc(&s->occ, occ_bytes) != hipSuccess || hipMalloc(&s->box, std::max<uint64_t>(1, nb) * 4) != hipSuccess ||
        (d->material_palette && hipMalloc(&s->palette, 256 * 3 * sizeof(float)) != hipSuccess))
        return cleanup(fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_create: device allocation failed"));
    if (d->material_palette &&
        hipMemcpy(s->palette, d->material_palette, 256 * 3 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(VX_ERR_HIP, "vx_scene_create: palette upload failed"));
    if (device) {
        s->nodes = const_cast<uint2*>(reinterpret_cast<const uint2*>(d->esvo_nodes));
        s->brick_data = const_cast<uint32_t*>(d->brick_data);
        s->bc1 = const_cast<uint2*>(reinterpret_cast<const uint2*>(d->bc1_colors));
        s->nrm = const_cast<uint4*>(reinterpret_cast<const uint4*>(d->dxt_normals));
    } else {
        if (hipMalloc(&s->nodes, node_bytes) != hipSuccess || hipMalloc(&s->brick_data, brick_bytes) != hipSuccess ||
            (d->bc1_colors && hipMalloc(&s->bc1, bc1_bytes) != hipSuccess) ||
            (d->dxt_normals && hipMalloc(&s->nrm, nrm_bytes) != hipSuccess))
            return cleanup(fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_create: device allocation failed"));
        if (hipMemcpy(s->nodes, d->esvo_nodes, node_bytes, hipMemcpyHostToDevice) != hipSuccess ||
            (nb && hipMemcpy(s->brick_data, d->brick_data, nb * 2048, hipMemcpyHostToDevice) != hipSuccess) ||
            (nb && d->bc1_colors && hipMemcpy(s->bc1, d->bc1_colors, nb * 256, hipMemcpyHostToDevice) != hipSuccess) ||
            (nb && d->dxt_normals && hipMemcpy(s->nrm, d->dxt_normals, nb * 512, hipMemcpyHostToDevice) != hipSuccess))
            return cleanup(fail(VX_ERR_HIP, "vx_scene_create: upload failed"));
    }
    if (nb) {
        const uint64_t words = nb * 16;
        k_build_occ<<<dim3((unsigned)((words + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0>>>(s->brick_data, words, s->occ);
        k_build_box<<<dim3((unsigned)((nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0>>>(s->occ, nb, s->box);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            return cleanup(fail(VX_ERR_HIP, "vx_scene_create: occupancy mask build failed"));
    }
    SceneView& v = s->view;
    v.nodes = s->nodes;
    v.brick_data = s->brick_data;
    v.occ = s->occ;
    v.box = s->box;
    v.bc1 = d->bc1_colors ? s->bc1 : nullptr;
    v.nrm = d->dxt_normals ? s->nrm : nullptr;
    v.palette = s->palette;
    v.node_count = static_cast<uint32_t>(d->node_count);
    v.brick_count = static_cast<uint32_t>(nb);
    const uint32_t N = B * 8;
    for (int i = 0; i < 3; ++i) v.gmin[i] = c.grid_min[i];
    v.Nf = static_cast<float>(N);
    v.k = v.Nf / sx;
    v.invN = 1.0f / v.Nf;
    v.Bf = static_cast<float>(B);
    v.min_scale = 22 - c.user_max_levels + 1;
    {
        // castRayImpl's worldRayLength = glm::length(worldMax - worldMin) and
        // worldSize.x (libraries/SVO/src/SVOTraversal.cpp:547-552)
        const float ex = c.grid_max[0] - c.grid_min[0], ey = c.grid_max[1] - c.grid_min[1],
                    ez = c.grid_max[2] - c.grid_min[2];
        v.lod_ray_length = std::sqrt((ex * ex + ey * ey) + ez * ez);
        v.lod_world_size = ex;
        v.lod_orig = 0.0f;
        v.lod_dir = 0.0f;
    }
    v.lds_off = 0;
    v.lds_count = 0;
    v.stack_depth = static_cast<uint32_t>(std::max(1, ilog2(B) - 1));
    v.stack_base = 0;
    v.aux_base = 0;
    s->lds_depth = 0;
    *out = s;
    return VX_OK;
}

// The scene's view for a plain launch of kBlock lanes (parks: the kernel
// postpones leaf visits); lds receives the dynamic LDS bytes.
static SceneView launch_view(const vx_scene* s, uint32_t& lds, bool parks)
{
    SceneView v = s->view;
    v.lds_off = 0;
    v.lds_count = staged_nodes(s, s->lds_depth);
    const uint32_t cap = (kMaxLdsBytes - kBlock * lane_lds_bytes(v.stack_depth, parks)) / 8u;
    if (v.lds_count > cap) v.lds_count = cap;
    lds = lay_out_lanes(v, v.lds_count, kBlock, parks);
    return v;
}

// Idle lanes that trigger a refill in k_trace_refill (measured on the 1080p
// cave: 16 beats 8, 24 and 32 on random and bounce rays).
constexpr uint32_t kRefillMin = 16;

// k_trace_refill pays off once each of its waves is dealt this many rays: with
// fewer (a call of under about a million rays) a wave has little to refill
// from, and the two launches' tails cost more than the refill saves, so the
// call stays one k_trace launch that postpones leaf visits in incoherent blocks.
constexpr uint64_t kSplitMinRange = 128;

// x with a * x = 1 mod m (a and m coprime, m < 2^32; 0 for m = 1).
static uint64_t inverse_mod(uint64_t a, uint64_t m)
{
    int64_t r0 = (int64_t)m, r1 = (int64_t)(a % m), t0 = 0, t1 = 1;
    while (r1 != 0) {
        const int64_t q = r0 / r1;
        std::swap(r0, r1);
        r1 -= q * r0;
        std::swap(t0, t1);
        t1 -= q * t0;
    }
    return (uint64_t)(t0 < 0 ? t0 + (int64_t)m : t0) % m;
}

// vx_trace on one stream: one k_trace launch for short calls, else two launches,
// k_trace<..., SPLIT> for the 64-ray blocks that share an origin and
// k_trace_refill for the others. Records are the same bits either way.
template <bool LOD>
static vx_status launch_trace_split(const vx_scene* s, const vx_lod* lod, const vx_ray* rays, uint64_t n, vx_hit* hits,
                                    const dim3& grid, uint32_t per, hipStream_t st)
{
    if ((n + 63u) >> 6 >= 0x100000000ull) return fail(VX_ERR_INVALID_ARG, "vx_trace: too many rays for one launch");
    // k_trace<..., SPLIT> never parks a leaf: its lane columns are two words shorter
    uint32_t ldsk = 0, lds = 0;
    SceneView vk = launch_view(s, ldsk, false), v = launch_view(s, lds, true);
    if (LOD) {
        vk.lod_orig = v.lod_orig = lod->ray_orig_size;
        vk.lod_dir = v.lod_dir = lod->ray_dir_size;
    }
    const bool staged = v.lds_count > 0;
    const dim3 block(kBlock);
    int per_cu = 0;
    if (staged) VX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_refill<true, LOD>, kBlock, lds));
    else VX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_refill<false, LOD>, kBlock, lds));
    if (per_cu < 1) return fail(VX_ERR_UNSUPPORTED, "vx_trace: the refill kernel does not fit one workgroup per CU");
    // as many workgroups as are resident at once, each wave dealt an equal range of at least 64 rays, and
    // fewer workgroups when the rays need fewer waves
    const uint64_t n_blocks = (n + 63u) >> 6;
    const uint64_t resident = (uint64_t)per_cu * (uint64_t)s->cu_count * (kBlock / 64);
    const uint64_t range = std::max<uint64_t>(64, (n_blocks * 64 + resident - 1) / resident);
    if (range < kSplitMinRange) {
        if (staged) k_trace<false, true, LOD><<<grid, block, lds, st>>>(v, rays, n, per, hits, nullptr, RefillDeal{0, 0});
        else k_trace<false, false, LOD><<<grid, block, lds, st>>>(v, rays, n, per, hits, nullptr, RefillDeal{0, 0});
        VX_HIP(hipGetLastError());
        return VX_OK;
    }
    // the block order's step: near n_blocks / golden ratio, odd and coprime to n_blocks
    uint64_t step = (uint64_t)((double)n_blocks * 0.6180339887498949) | 1u;
    while (std::gcd(step, n_blocks) != 1u) step += 2u;
    const RefillDeal deal{range, inverse_mod(step % n_blocks, n_blocks)};
    if (vk.lds_count > 0) k_trace<false, true, LOD, true><<<grid, block, ldsk, st>>>(vk, rays, n, per, hits, nullptr, deal);
    else k_trace<false, false, LOD, true><<<grid, block, ldsk, st>>>(vk, rays, n, per, hits, nullptr, deal);
    VX_HIP(hipGetLastError());
    const uint64_t waves = (n_blocks * 64 + range - 1) / range;
    const dim3 rgrid(static_cast<unsigned>((waves + kBlock / 64 - 1) / (kBlock / 64)));
    if (staged) k_trace_refill<true, LOD><<<rgrid, block, lds, st>>>(v, rays, n, range, step % n_blocks, kRefillMin, hits);
    else k_trace_refill<false, LOD><<<rgrid, block, lds, st>>>(v, rays, n, range, step % n_blocks, kRefillMin, hits);
    VX_HIP(hipGetLastError());
    return VX_OK;
}

static vx_status launch_trace(vx_scene* s, const vx_ray* rays, uint64_t n, vx_hit* hits, uint64_t* counters,
                              hipStream_t st, const vx_lod* lod = nullptr)
{
    if (!s || (n && (!rays || !hits))) return fail(VX_ERR_INVALID_ARG, "vx_trace: null argument");
    if (vx_status st = scene_usable(s, "vx_trace"); st != VX_OK) return st;
    if (n == 0) return VX_OK;
    VX_DEVICE_GUARD(s->device);
    const uint64_t chunks = (n + kBlock - 1) / kBlock;
    const uint64_t per = (chunks + kXcds - 1) / kXcds;
    if (per * kXcds > 0x7FFFFFFFull) return fail(VX_ERR_INVALID_ARG, "vx_trace: too many rays for one launch");
    const dim3 grid(static_cast<unsigned>(per * kXcds)), block(kBlock);
    uint32_t lds = 0;
    const SceneView v = launch_view(s, lds, true); // the counting k_trace postpones leaf visits
    const bool staged = v.lds_count > 0;
    if (lod && lod->ray_dir_size > 0.0f) { // LODParameters::isEnabled
        return launch_trace_split<true>(s, lod, rays, n, hits, grid, static_cast<uint32_t>(per), st);
    } else if (counters) {
        VX_HIP(hipMemsetAsync(counters, 0, 8 * sizeof(uint64_t), st));
        auto* c = reinterpret_cast<unsigned long long*>(counters);
        if (staged) k_trace<true, true><<<grid, block, lds, st>>>(v, rays, n, static_cast<uint32_t>(per), hits, c, RefillDeal{0, 0});
        else k_trace<true, false><<<grid, block, lds, st>>>(v, rays, n, static_cast<uint32_t>(per), hits, c, RefillDeal{0, 0});
    } else {
        return launch_trace_split<false>(s, nullptr, rays, n, hits, grid, static_cast<uint32_t>(per), st);
    }
    VX_HIP(hipGetLastError());
    return VX_OK;
}

} // namespace vx

extern "C" {

const char* vx_last_error(void) { return vx::g_last_error.c_str(); }

const char* vx_version(void) { return "vixen-hip 0.2 gfx950 wave64 strict-fp32"; }

vx_status vx_ctx_create(int device, vx_ctx** out)
{
    if (!out) return vx::fail(VX_ERR_INVALID_ARG, "vx_ctx_create: null out");
    *out = nullptr;
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess)
        return vx::fail(VX_ERR_HIP, std::string("vx_ctx_create: hipGetDeviceCount failed: ") + hipGetErrorString(e));
    if (device < 0 || device >= count)
        return vx::fail(VX_ERR_INVALID_ARG, "vx_ctx_create: device " + std::to_string(device) + " not in [0, " +
                                                std::to_string(count) + ")");
    if (hipSetDevice(device) != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_ctx_create: hipSetDevice failed");
    vx::scratch_init(device); // vx_albedo_backward_noalloc's workspace on this device
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
    auto* c = new (std::nothrow) vx_ctx{device, cus};
    if (!c) return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_ctx_create: out of memory");
    *out = c;
    return VX_OK;
}

vx_status vx_ctx_destroy(vx_ctx* ctx)
{
    delete ctx;
    return VX_OK;
}

vx_status vx_scene_create(vx_ctx* ctx, const vx_scene_desc* desc, vx_scene** out)
{
    try {
        return vx::scene_create(ctx, desc, out, false);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_create: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_create: unexpected exception");
    }
}

vx_status vx_scene_create_device(vx_ctx* ctx, const vx_scene_desc* desc, vx_scene** out)
{
    try {
        return vx::scene_create(ctx, desc, out, true);
    } catch (const std::bad_alloc&) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_create_device: out of memory");
    } catch (...) {
        return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_create_device: unexpected exception");
    }
}

vx_status vx_scene_validate(const vx_scene_desc* desc)
{
    try {
        uint32_t level_prefix[4];
        return vx::scene_validate_impl(desc, level_prefix);
    } catch (...) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_scene_validate: unexpected exception");
    }
}

vx_status vx_scene_retain(vx_scene* s)
{
    if (!s) return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_retain: null scene");
    s->refs.fetch_add(1, std::memory_order_relaxed);
    return VX_OK;
}

vx_status vx_scene_destroy(vx_scene* s)
{
    if (!s) return VX_OK;
    if (s->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return VX_OK; // other references remain
    const vx::DeviceGuard guard(s->device); // frees proceed even if it fails
    vx_albedo_state_free(s->albedo);
    if (!s->borrowed) {
        (void)hipFree(s->nodes);
        (void)hipFree(s->brick_data);
        (void)hipFree(s->bc1);
        (void)hipFree(s->nrm);
    }
    (void)hipFree(s->occ);
    (void)hipFree(s->box);
    (void)hipFree(s->palette);
    delete s;
    return VX_OK;
}

vx_status vx_scene_info(const vx_scene* s, uint64_t* node_count, uint64_t* brick_count, int32_t* user_max_levels,
                        int32_t* bricks_per_axis)
{
    if (!s) return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_info: null scene");
    if (node_count) *node_count = s->node_count;
    if (brick_count) *brick_count = s->brick_count;
    if (user_max_levels) *user_max_levels = s->cfg.user_max_levels;
    if (bricks_per_axis) *bricks_per_axis = s->cfg.bricks_per_axis;
    return VX_OK;
}

vx_status vx_scene_revalidate(vx_scene* s)
{
    if (!s) return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_revalidate: null scene");
    VX_DEVICE_GUARD(s->device);
    std::vector<uint32_t> host(s->node_count * 2);
    if (hipMemcpy(host.data(), s->nodes, s->node_count * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return vx::fail(VX_ERR_HIP, "vx_scene_revalidate: reading the descriptors failed");
    vx_scene_desc hd{};
    hd.esvo_nodes = host.data();
    hd.node_count = s->node_count;
    hd.brick_data = s->brick_data;
    hd.brick_count = s->brick_count;
    hd.config = &s->cfg;
    uint32_t level_prefix[4];
    if (vx_status st = vx::scene_validate_impl(&hd, level_prefix); st != VX_OK) {
        s->invalid.store(true, std::memory_order_release);
        return st;
    }
    const uint64_t nb = s->brick_count;
    if (nb) {
        const uint64_t words = nb * 16;
        vx::k_build_occ<<<dim3((unsigned)((words + vx::kBlock - 1) / vx::kBlock)), dim3(vx::kBlock), 0, 0>>>(
            s->brick_data, words, s->occ);
        vx::k_build_box<<<dim3((unsigned)((nb + vx::kBlock - 1) / vx::kBlock)), dim3(vx::kBlock), 0, 0>>>(s->occ, nb,
                                                                                                           s->box);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            s->invalid.store(true, std::memory_order_release);
            return vx::fail(VX_ERR_HIP, "vx_scene_revalidate: occupancy mask build failed");
        }
    }
    std::memcpy(s->level_prefix, level_prefix, sizeof level_prefix);
    // the differentiable renderer's voxel rows come from the occupancy: rebuilt on next use
    vx_albedo_state_free(s->albedo);
    s->albedo = nullptr;
    s->invalid.store(false, std::memory_order_release);
    return VX_OK;
}

vx_status vx_scene_set_lds_depth(vx_scene* s, uint32_t depth)
{
    if (!s || depth > 4) return vx::fail(VX_ERR_INVALID_ARG, "vx_scene_set_lds_depth: depth must be 0..4");
    s->lds_depth = depth;
    return VX_OK;
}

vx_status vx_trace(vx_scene* s, const vx_ray* rays, uint64_t n, vx_hit* hits, vx_stream stream)
{
    return vx::launch_trace(s, rays, n, hits, nullptr, static_cast<hipStream_t>(stream));
}

vx_status vx_trace_lod(vx_scene* s, const vx_ray* rays, uint64_t n, const vx_lod* lod, vx_hit* hits,
                       vx_stream stream)
{
    if (!lod) return vx::fail(VX_ERR_INVALID_ARG, "vx_trace_lod: null lod");
    return vx::launch_trace(s, rays, n, hits, nullptr, static_cast<hipStream_t>(stream), lod);
}

// LODParameters::fromCamera (libraries/SVO/include/SVOLOD.h:83-90)
vx_lod vx_lod_from_camera(float fov_y, int32_t screen_height)
{
    const float pixel_angle = fov_y / static_cast<float>(screen_height);
    return vx_lod{0.0f, 2.0f * std::tan(pixel_angle * 0.5f)};
}

// LODParameters::fromCameraWithNearPlane (SVOLOD.h:105-116)
vx_lod vx_lod_from_camera_near(float fov_y, int32_t screen_height, float near_plane)
{
    const float dir = vx_lod_from_camera(fov_y, screen_height).ray_dir_size;
    return vx_lod{near_plane * dir, dir};
}

// LODParameters::withBias (SVOLOD.h:169-175)
vx_lod vx_lod_with_bias(vx_lod lod, float bias)
{
    const float m = std::pow(2.0f, bias);
    return vx_lod{lod.ray_orig_size * m, lod.ray_dir_size * m};
}

vx_status vx_trace_counters(vx_scene* s, const vx_ray* rays, uint64_t n, vx_hit* hits, uint64_t* counters,
                            vx_stream stream)
{
    if (!counters) return vx::fail(VX_ERR_INVALID_ARG, "vx_trace_counters: null counters");
    return vx::launch_trace(s, rays, n, hits, counters, static_cast<hipStream_t>(stream));
}

vx_status vx_render(vx_scene* s, const vx_camera* cam, uint32_t w, uint32_t h, float tmax, vx_hit* hits,
                    uint32_t* rgba8, vx_stream stream)
{
    if (!s || !cam || (!hits && !rgba8)) return vx::fail(VX_ERR_INVALID_ARG, "vx_render: null argument");
    if (w == 0 || h == 0) return VX_OK;
    if (w > 65536 || h > 65536) return vx::fail(VX_ERR_INVALID_ARG, "vx_render: image too large");
    if (vx_status st = vx::scene_usable(s, "vx_render"); st != VX_OK) return st;
    VX_DEVICE_GUARD(s->device);
    const vx::CamParams c = vx::make_cam(cam);
    const uint32_t tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16, ntiles = tiles_x * tiles_y;
    const uint32_t per = (ntiles + vx::kXcds - 1) / vx::kXcds;
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t lds = 0;
    const vx::SceneView v = vx::launch_view(s, lds, false); // k_render visits leaves at once
    const dim3 grid(per * vx::kXcds), block(vx::kBlock);
    const int shade = !rgba8 ? vx::SHADE_NONE : ((v.nrm && v.bc1) ? vx::SHADE_COMPRESSED : vx::SHADE_FACE);
    auto launch = [&](auto kernel) {
        kernel<<<grid, block, lds, st>>>(v, c, w, h, tmax, tiles_x, ntiles, per, hits, rgba8);
    };
    if (v.lds_count > 0) {
        if (shade == vx::SHADE_NONE) launch(vx::k_render<true, vx::SHADE_NONE>);
        else if (shade == vx::SHADE_FACE) launch(vx::k_render<true, vx::SHADE_FACE>);
        else launch(vx::k_render_compressed<true>);
    } else {
        if (shade == vx::SHADE_NONE) launch(vx::k_render<false, vx::SHADE_NONE>);
        else if (shade == vx::SHADE_FACE) launch(vx::k_render<false, vx::SHADE_FACE>);
        else launch(vx::k_render_compressed<false>);
    }
    if (hipGetLastError() != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_render: launch failed");
    return VX_OK;
}

vx_status vx_bounce_rays(const vx_scene* s, const vx_ray* primary, const vx_hit* hits, uint64_t n, uint32_t seed,
                         vx_ray* out, vx_stream stream)
{
    if (!s || (n && (!primary || !hits || !out))) return vx::fail(VX_ERR_INVALID_ARG, "vx_bounce_rays: null argument");
    if (n == 0) return VX_OK;
    if (n > 0xFFFFFFFFull * vx::kBlock) return vx::fail(VX_ERR_INVALID_ARG, "vx_bounce_rays: too many rays");
    VX_DEVICE_GUARD(s->device);
    vx::k_bounce_rays<<<dim3((unsigned)((n + vx::kBlock - 1) / vx::kBlock)), dim3(vx::kBlock), 0,
                        static_cast<hipStream_t>(stream)>>>(primary, hits, n, seed, s->view.k, out);
    if (hipGetLastError() != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_bounce_rays: launch failed");
    return VX_OK;
}

vx_status vx_camera_rays(const vx_camera* cam, uint32_t w, uint32_t h, float tmax, vx_ray* rays, vx_stream stream)
{
    if (!cam || !rays) return vx::fail(VX_ERR_INVALID_ARG, "vx_camera_rays: null argument");
    const uint64_t n = (uint64_t)w * h;
    if (n == 0) return VX_OK;
    if (w > 65536 || h > 65536) return vx::fail(VX_ERR_INVALID_ARG, "vx_camera_rays: image too large");
    const vx::CamParams c = vx::make_cam(cam);
    vx::k_camera_rays<<<dim3((unsigned)((n + vx::kBlock - 1) / vx::kBlock)), dim3(vx::kBlock), 0,
                        static_cast<hipStream_t>(stream)>>>(c, w, h, tmax, rays);
    if (hipGetLastError() != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_camera_rays: launch failed");
    return VX_OK;
}

} // extern "C"
