// vx_trace.hpp -- device-side traversal for gfx950: ESVO -> brick DDA -> colour.
//
// One ray per lane of a wave64. The arithmetic is strict FP32 (the library is
// compiled with -ffp-contract=off and correctly rounded divide/sqrt) and keeps
// the evaluation order of the reference traversal so results are bit-identical
// to a strict-FP32 CPU evaluation of the same steps:
//   ray set-up / volume entry : libraries/SVO/src/SVOTraversal.cpp:117-149,238-259,516-583
//   coefficients              : libraries/SVO/src/SVOTraversal.cpp:699-728
//   init/PUSH/ADVANCE/POP     : shaders/ESVOTraversal.glsl:84-372
//   loop + t_max restore      : shaders/VoxelRayMarch_Compressed.comp:396-478
//   leaf descriptor           : shaders/VoxelRayMarch_Compressed.comp:291-305
//   brick cell from ESVO pos  : libraries/SVO/src/SVOBrickDDA.cpp:85-93
//   brick DDA, t, normal      : libraries/SVO/src/SVOBrickDDA.cpp:187-225,390-533
//   BC1 decode                : libraries/VoxelData/src/Compression/DXT1Compressor.cpp:14-36
//
// Layout choices that are this backend's own:
//   - the traversal stack only holds the scales the scene's tree can reach
//     (22 down to one above the brick level: log2(bricks per axis) - 1 entries,
//     at most VX_STACK) and lives in LDS, in a contiguous column per lane
//     (columns an odd number of words apart: bank-conflict free, no VGPRs, no
//     scratch); every entry a ray reads was written earlier on its own path,
//     so the stack needs no initialisation;
//   - the ray's brick-DDA inputs (local origin and direction, reciprocal
//     direction, traced segment: 11 words) and, where visits are postponed, a
//     pending leaf (2 words) are needed only at brick visits, so they too live
//     in the lane's LDS column instead of VGPRs across the ESVO loop (more
//     waves per SIMD); a finished ray's record goes there as well, so the loop
//     carries no hit-record registers;
//   - a per-brick box of occupied voxels, derived at upload, rejects brick
//     segments that cannot hit before the DDA is set up;
//   - for incoherent rays, brick visits are postponed until most live lanes of
//     the wave wait at one, so the wave runs the DDA loop once for all of them
//     (trace_ray);
//   - brick occupancy is read from a 64-byte-per-brick bit mask derived from
//     brickData at upload (bit lin of the brick's 16 words), so a brick's DDA
//     touches one cache line instead of 2 KiB; the hit voxel's brickData word
//     is read only when the scene has no BC1 colours;
//   - the brick DDA is software-pipelined: the voxel sequence does not depend
//     on occupancy, so each step computes the next voxel and issues its mask
//     load before testing the current voxel (one load in flight per lane);
//     the sequence of voxels tested, and so the result, is unchanged.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vixen_hip.h"

namespace vx {

constexpr int VX_STACK = 8;          // supports up to 9 levels above the bricks (N <= 4096)
constexpr int ESVO_MAX_SCALE = 22;
constexpr int MAX_ITERS = 512;       // shaders/ESVOTraversal.glsl:25
constexpr float T_EPS = 1e-6f;       // shaders/ESVOTraversal.glsl:26
constexpr float DIR_EPS = 1e-5f;     // shaders/ESVOTraversal.glsl:27, SVOTraversal.cpp:707

struct SceneView {
    const uint2* __restrict__ nodes;
    const uint32_t* __restrict__ brick_data;  // 512 words per brick (colour when no BC1)
    const uint32_t* __restrict__ occ;         // 16 words per brick: occupancy bits
    // one word per brick: the box of its occupied voxels, lo x|y|z in bits 0-8
    // and hi x|y|z in bits 9-17, 3 bits each (the whole brick when it is empty)
    const uint32_t* __restrict__ box;
    const uint2* __restrict__ bc1;            // 32 blocks per brick or nullptr
    const uint4* __restrict__ nrm;            // 32 DXT normal blocks per brick or nullptr
    const float* __restrict__ palette;        // 256 x RGB by material id (brick word & 0xFF) or nullptr
    uint32_t node_count;
    uint32_t brick_count;
    float gmin[3];
    float k;          // N / (gmax.x - gmin.x): world -> voxel units
    float Nf;         // N
    float invN;       // 1 / N (a power of two)
    float Bf;         // bricks per axis
    int min_scale;
    uint32_t lds_off;    // first LDS slot of this scene's staged descriptors
    uint32_t lds_count;  // descriptors [0, lds_count) are read from LDS
    uint32_t stack_base; // first LDS word of the per-lane columns (stack, then aux words)
    uint32_t stack_depth; // stack entries per lane (levels above the bricks - 1, >= 1)
    uint32_t aux_base;   // offset of the brick-DDA inputs in a lane's column
    uint32_t col_words;  // words per lane column (odd)
    // screen-space LOD (LOD kernels only): ray cone size at the origin and
    // growth per unit distance, |grid_max - grid_min|, and the grid's x extent
    float lod_orig, lod_dir, lod_ray_length, lod_world_size;
};

struct TraceCounts {
    uint32_t iters, leaves, voxels, hits;
};

__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }
// One-instruction min/max (IEEE 754-2019 minimum/maximum: v_minimum3_f32 /
// v_maximum3_f32, no operand quieting needed). They equal mn/mx whenever the
// operands are not NaN and not a (-0, +0) pair; every use states why that holds.
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t popc8(uint32_t v) { return __builtin_popcount(v & 0xFFu); }

// Scene arrays always live in device global memory. Loads go through a
// global-address-space pointer, so they compile to global loads (base in SGPRs,
// vector memory counter only) even where the compiler cannot see where a
// pointer came from: the batched kernel reads its scene views from a table in
// memory, and with generic pointers it emitted flat loads, which also count
// against the LDS wait counter and so made every wait on the traversal's LDS
// columns wait for outstanding memory loads too.
template <class T>
using gmem_ptr = const __attribute__((address_space(1))) T*;

// A 4- or 8-byte scene element loaded as a plain integer through a global
// pointer (a vector type's copy constructor would take the address back as a
// generic reference and the load would become a flat load again).
template <class T>
__device__ __forceinline__ T ld_global(const char* p)
{
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, *(gmem_ptr<unsigned long long>)p);
    } else {
        static_assert(sizeof(T) == 4, "4- or 8-byte scene elements");
        return __builtin_bit_cast(T, *(gmem_ptr<uint32_t>)p);
    }
}

// Element i of a scene array through a 32-bit byte offset from its (uniform)
// base, so the load takes the base in SGPRs and only a 32-bit offset per lane,
// with no 64-bit address arithmetic. Valid while i * sizeof(T) < 2^32: the
// scene checks bound the descriptor count below 2^29, and 24-bit brick ids keep
// the occupancy, box and BC1 pools below 4 GiB.
template <class T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, uint32_t i)
{
    return ld_global<T>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(T));
}
// The same with a byte offset.
template <class T>
__device__ __forceinline__ T ldb32(const T* __restrict__ base, uint32_t byte_offset)
{
    return ld_global<T>(reinterpret_cast<const char*>(base) + byte_offset);
}

__device__ __forceinline__ void store_miss(vx_hit& h, uint32_t iters)
{
    h.flags = iters << 16;
    h.t = __builtin_inff();
    h.leaf_node = 0xFFFFFFFFu;
    h.brick = 0xFFFFFFFFu;
    h.voxel = 0xFFFFFFFFu;
    h.rgb[0] = h.rgb[1] = h.rgb[2] = 0.0f;
}

__device__ __forceinline__ void load_ray(const vx_ray* __restrict__ rays, uint64_t i, float o[3], float d[3],
                                         float& tmin, float& tmax)
{
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
    tmin = a.w;
    d[0] = b.x;
    d[1] = b.y;
    d[2] = b.z;
    tmax = b.w;
}

__device__ __forceinline__ void store_hit(vx_hit* __restrict__ out, uint64_t i, const vx_hit& h)
{
    // written once and never read back by the kernel: streamed (nt) stores
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    u32x4* dst = reinterpret_cast<u32x4*>(out + i);
    __builtin_nontemporal_store(u32x4{h.flags, __float_as_uint(h.t), h.leaf_node, h.brick}, dst);
    __builtin_nontemporal_store(u32x4{h.voxel, __float_as_uint(h.rgb[0]), __float_as_uint(h.rgb[1]),
                                      __float_as_uint(h.rgb[2])}, dst + 1);
}

// BC1 decode of one texel (integer extraction, u32->f32 round-to-nearest, x 2^-32,
// blend with fl(2/3), fl(1/3), no contraction).
__device__ __forceinline__ void bc1_decode(uint2 blk, uint32_t texel, float rgb[3])
{
    const uint32_t head = blk.x, bits = blk.y;
    const float S = 0x1p-32f;
    const uint32_t sel = (bits >> (2u * texel)) & 3u;
    const float r0 = (float)(head << 16) * S, g0 = (float)(head << 21) * S, b0 = (float)(head << 27) * S;
    const float r1 = (float)head * S, g1 = (float)(head << 5) * S, b1 = (float)(head << 11) * S;
    const float c23 = 2.0f / 3.0f, c13 = 1.0f / 3.0f;
    if (sel == 0u) {
        rgb[0] = r0; rgb[1] = g0; rgb[2] = b0;
    } else if (sel == 1u) {
        rgb[0] = r1; rgb[1] = g1; rgb[2] = b1;
    } else if (sel == 2u) {
        rgb[0] = r0 * c23 + r1 * c13;
        rgb[1] = g0 * c23 + g1 * c13;
        rgb[2] = b0 * c23 + b1 * c13;
    } else {
        rgb[0] = r0 * c13 + r1 * c23;
        rgb[1] = g0 * c13 + g1 * c23;
        rgb[2] = b0 * c13 + b1 * c23;
    }
}

__device__ __forceinline__ void rgba8_decode(uint32_t w, float rgb[3])
{
    rgb[0] = (float)(w & 255u) / 255.0f;
    rgb[1] = (float)((w >> 8) & 255u) / 255.0f;
    rgb[2] = (float)((w >> 16) & 255u) / 255.0f;
}

// The uncompressed pipeline's colour of a voxel word: getMaterialColor(word & 0xFF)
// (shaders/VoxelRayMarch.comp:189-191) as a 256-entry palette.
__device__ __forceinline__ void palette_decode(const float* __restrict__ palette, uint32_t w, float rgb[3])
{
    const uint32_t c = (w & 255u) * 3u;
    rgb[0] = ld32(palette, c);
    rgb[1] = ld32(palette, c + 1u);
    rgb[2] = ld32(palette, c + 2u);
}

// Ray set-up in the grid's local voxel space [0,N]^3. Returns false on a miss
// or an invalid ray.
__device__ __forceinline__ bool setup_ray(const SceneView& s, const float o[3], const float d[3], float tmin,
                                          float tmax, float ol[3], float dl[3], float du[3], float& tstart,
                                          float& texit, float invd[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (!__builtin_isfinite(o[i]) || !__builtin_isfinite(d[i])) return false;
    const float len2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!(len2 > 0.0f) || !__builtin_isfinite(len2)) return false;
    const float inv = 1.0f / __builtin_sqrtf(len2);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        du[i] = d[i] * inv;
        if (!__builtin_isfinite(du[i])) return false;
        ol[i] = (o[i] - s.gmin[i]) * s.k;
        dl[i] = du[i] * s.k;
        if (__builtin_fabsf(dl[i]) < 1e-8f) {
            if (ol[i] < 0.0f || ol[i] > s.Nf) return false;
            invd[i] = (dl[i] >= 0.0f) ? 1e20f : -1e20f;
        } else {
            invd[i] = 1.0f / dl[i];
        }
    }
    float tn[3], tf[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float t0 = (0.0f - ol[i]) * invd[i];
        const float t1 = (s.Nf - ol[i]) * invd[i];
        // t0, t1: products of finite values (never NaN) and never both zero, so
        // IEEE min/max equal the reference's ternaries here
        tn[i] = fmn(t0, t1);
        tf[i] = fmx(t0, t1);
    }
    float tEntry = mx(mx(tn[0], tn[1]), tn[2]);
    float tExit = mn(mn(tf[0], tf[1]), tf[2]);
    if (!(tEntry <= tExit && tExit >= 0.0f)) return false;
    tEntry = mx(tEntry, tmin);
    tExit = mn(tExit, tmax);
    if (tEntry >= tExit || tExit < 0.0f) return false;
    tstart = mx(tEntry, 0.0f);
    texit = tExit;
    return true;
}

// Brick DDA in local voxel space. Returns true on a hit, with the hit voxel
// (x + 8y + 64z) and the axis of the step into it (1 x, 2 y, 3 z; 0 when it is
// the first voxel tested); hit_entry then gives the record's t and face.
// rinv[i] is the ray's brick reciprocal, 1/dl[i] (+-1e8 when |dl[i]| < 1e-8),
// computed once per ray: the reference's per-brick 1/dl and 1/|dl| are the
// same correctly rounded values (RN(1/|x|) = |RN(1/x)|).
template <bool COUNT>
__device__ __forceinline__ bool brick_dda(const SceneView& s, uint32_t brick, const int bc[3], const float ol[3],
                                          const float dl[3], const float rinv[3], float tstart, float texit,
                                          uint32_t& voxel, uint32_t& step_axis, TraceCounts& cnt)
{
    float bmin[3], tn[3], tf[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        bmin[i] = (float)(bc[i] * 8);
        const float bmax = bmin[i] + 8.0f;
        const float t0 = (bmin[i] - ol[i]) * rinv[i];
        const float t1 = (bmax - ol[i]) * rinv[i];
        tn[i] = fmn(t0, t1); // as in setup_ray: t0, t1 never NaN, never both zero
        tf[i] = fmx(t0, t1);
    }
    float btmin = mx(mx(tn[0], tn[1]), tn[2]);
    float btmax = mn(mn(tf[0], tf[1]), tf[2]);
    btmin = mx(btmin, tstart);
    btmax = mn(btmax, texit);
    if (btmin > btmax) return false;
    if (!COUNT) {
        // Occupied-box pretest. Every voxel the DDA below would test is crossed
        // by the segment [btmin, btmax] to within its rounding error (under 0.02
        // voxel for grids up to 4096^3), so a segment that misses the brick's
        // occupied voxels' box, grown by 0.25 voxel, finds no hit: return what
        // the DDA would. (The counting variant keeps the full walk, so its
        // counts are the reference algorithm's.)
        const uint32_t b = ld32(s.box, brick);
        float en = btmin, ex = btmax;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float lo = bmin[i] + ((float)((b >> (3 * i)) & 7u) - 0.25f);
            const float hi = bmin[i] + ((float)((b >> (9 + 3 * i)) & 7u) + 1.25f);
            const float t0 = (lo - ol[i]) * rinv[i];
            const float t1 = (hi - ol[i]) * rinv[i];
            // finite operands; the result only feeds en > ex, where -0 and +0 agree
            en = fmx(en, fmn(t0, t1));
            ex = fmn(ex, fmx(t0, t1));
        }
        if (en > ex) return false;
    }

    int cur[3], step[3];
    float tDelta[3], tNext[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float entry = ol[i] + dl[i] * btmin;
        const int c = (int)__builtin_floorf(entry - bmin[i]);
        cur[i] = c < 0 ? 0 : (c > 7 ? 7 : c);
        // positive: (bmin + cur + 1 - entry) / dl; negative: (entry - (bmin + cur)) / |dl|;
        // one division by |dl| either way, no divergent branch
        const bool pos = dl[i] > 0.0f;
        const bool flat = __builtin_fabsf(dl[i]) < 1e-8f;
        const float nb = bmin[i] + (float)(cur[i] + (pos ? 1 : 0));
        const float q = (pos ? nb - entry : entry - nb) / __builtin_fabsf(dl[i]);
        step[i] = flat ? 0 : (pos ? 1 : -1);
        tDelta[i] = flat ? 3.402823466e+38f : __builtin_fabsf(rinv[i]);
        tNext[i] = flat ? 3.402823466e+38f : btmin + q;
    }
    // cur starts inside the brick (clamped above). Each pass steps the DDA first
    // (the reference steps after the test: the same voxel order, the same exits at
    // t > btmax, outside the brick or after 24 voxels) and loads the next voxel's
    // mask word, always from inside this brick, before it tests the current one.
    // The step is written with selects (no divergent branches); the hit record is
    // evaluated once, after the loop.
    //
    // The voxel is tracked as its linear index alone, lin = x + 8y + 64z. A step
    // adds +-1, +-8 or +-64, which leaves the bits below the stepped axis's 3-bit
    // field alone; it stays in the brick exactly when nothing above the field
    // changes either (a coordinate of 8 carries out of the field, one of -1
    // borrows), i.e. when lin ^ nlin < 8, 64 or 512 for an x, y or z step: the
    // reference's 0 <= x, y, z <= 7 test on the new coordinates.
    const uint32_t occ = brick * 64u; // byte offset of the brick's 16 mask words (64-byte aligned)
    const uint32_t d0 = (uint32_t)step[0], d1 = (uint32_t)(8 * step[1]), d2 = (uint32_t)(64 * step[2]);
    uint32_t lin = (uint32_t)(cur[0] + 8 * cur[1] + 64 * cur[2]);
    // mask word of voxel v: byte occ + 4 * (v >> 5) = occ | ((v >> 3) & 60) (v < 512)
    uint32_t word = ldb32(s.occ, occ | ((lin >> 3) & 60u));
    bool hit = false;
    uint32_t last = 0u; // axis of the step into the current voxel (1 x, 2 y, 3 z; 0 = entry voxel)
    for (int it = 0;; ++it) {
        if (COUNT) cnt.voxels += 1;
        // bitwise & on the conditions: no short-circuit branches
        const bool ax = (tNext[0] < tNext[1]) & (tNext[0] < tNext[2]);
        const bool ay = !ax & (tNext[1] < tNext[2]);
        const bool az = !ax & !ay;
        const float tsel = ax ? tNext[0] : (ay ? tNext[1] : tNext[2]);
        const uint32_t nlin = lin + (ax ? d0 : (ay ? d1 : d2));
        const bool inside = (nlin ^ lin) < (ax ? 8u : (ay ? 64u : 512u));
        const bool more = !(tsel > btmax) & (it + 1 < 24) & inside;
        const uint32_t nword = ldb32(s.occ, occ | ((nlin >> 3) & 60u)); // in this brick even when nlin is not
        const bool occupied = __builtin_amdgcn_ubfe(word, lin, 1u) != 0u; // bit lin mod 32
        if (occupied | !more) {
            hit = occupied;
            break;
        }
        const float s0 = tNext[0] + tDelta[0], s1 = tNext[1] + tDelta[1], s2 = tNext[2] + tDelta[2];
        tNext[0] = ax ? s0 : tNext[0];
        tNext[1] = ay ? s1 : tNext[1];
        tNext[2] = az ? s2 : tNext[2];
        lin = nlin;
        word = nword;
        last = ax ? 1u : (ay ? 2u : 3u);
    }
    if (!hit) return false;
    voxel = lin;
    step_axis = last;
    return true;
}

// The hit record's entry time and face for hit voxel `voxel` of the brick at
// grid cell bc (SVOBrickDDA.cpp:470-502), flags = hit | axis << 1 | sign << 3
// | step axis << 4.
__device__ __forceinline__ void hit_entry(const int bc[3], uint32_t voxel, uint32_t step_axis, const float ol[3],
                                          const float dl[3], float tstart, uint32_t& flags, float& t)
{
    const int cur[3] = {(int)(voxel & 7u), (int)((voxel >> 3) & 7u), (int)(voxel >> 6)};
    // Entry time of the hit voxel per axis: the reference takes min(t0, t1) of
    // its two faces; rounding is monotone, so that is the near face's quotient
    // ((vmin - o)/d for d > 0, (vmax - o)/d for d < 0): one division per axis.
    float tnh[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float vmin = (float)(bc[i] * 8) + (float)cur[i];
        const float vmax = vmin + 1.0f;
        const float tq = ((dl[i] > 0.0f ? vmin : vmax) - ol[i]) / dl[i];
        tnh[i] = __builtin_fabsf(dl[i]) < 1e-8f ? -__builtin_inff() : tq;
    }
    float tHit = mx(mx(tnh[0], tnh[1]), tnh[2]);
    tHit = mx(tHit, tstart);
    // sign bits computed for all axes and selected as values: an indexed
    // read of dl[] here would make the ray state addressable (scratch)
    const uint32_t sx = dl[0] > 0.0f ? 0u : 1u, sy = dl[1] > 0.0f ? 0u : 1u, sz = dl[2] > 0.0f ? 0u : 1u;
    uint32_t axis, signpos;
    if (tnh[0] >= tnh[1] && tnh[0] >= tnh[2]) {
        axis = 0u;
        signpos = sx;
    } else if (tnh[1] >= tnh[2]) {
        axis = 1u;
        signpos = sy;
    } else {
        axis = 2u;
        signpos = sz;
    }
    flags = 1u | (axis << 1) | (signpos << 3) | (step_axis << 4);
    t = tHit;
}

// Descriptor fetch: a prefix of the descriptor array (the top levels of a
// breadth-first tree) may be staged in LDS by the launching kernel.
// The launching kernel's dynamic LDS: staged descriptors first, then one column
// of col_words consecutive words per lane: the traversal stack (stack_depth
// parent pointers, then stack_depth t_max words) and the kAuxWords brick-DDA
// inputs. Columns are an odd number of words apart, so the lanes of a wave
// reading word j of their columns hit distinct banks, and every word of a
// column is one base address plus an immediate offset.
extern __shared__ uint2 vx_lds_nodes[];

// STAGED is a compile-time choice: without staging the fetch is a plain global
// load (a runtime LDS-or-global choice compiles to a slower flat load).
template <bool STAGED>
__device__ __forceinline__ uint2 fetch_node(const SceneView& s, uint32_t i)
{
    if (STAGED && i < s.lds_count) return vx_lds_nodes[s.lds_off + i];
    return ld32(s.nodes, i);
}

__device__ __forceinline__ uint32_t* lane_word(const SceneView& s, uint32_t base, uint32_t j)
{
    return reinterpret_cast<uint32_t*>(vx_lds_nodes) + s.stack_base + threadIdx.x * s.col_words + base + j;
}
__device__ __forceinline__ uint32_t* stack_word(const SceneView& s, int k, bool tmax)
{
    return lane_word(s, 0u, (tmax ? s.stack_depth : 0u) + (uint32_t)k);
}
__device__ __forceinline__ void stack_put(const SceneView& s, int k, uint32_t ptr, float t)
{
    *stack_word(s, k, false) = ptr;
    *stack_word(s, k, true) = f2u(t);
}
__device__ __forceinline__ float stack_tmax(const SceneView& s, int k) { return u2f(*stack_word(s, k, true)); }
__device__ __forceinline__ uint32_t stack_ptr(const SceneView& s, int k) { return *stack_word(s, k, false); }

// A finished ray's record is written to words 0-4 of its lane's LDS column
// (over the ray inputs, dead by then) and assembled into a vx_hit once the
// traversal loop is left, so the loop carries no hit-record registers:
//   word 0 flags (a miss: iterations << 16 only), 1 t, 2 leaf descriptor,
//   3 brick, 4 voxel.
__device__ __forceinline__ void rec_miss(const SceneView& s, uint32_t iters)
{
    *lane_word(s, s.aux_base, 0) = iters << 16;
}
__device__ __forceinline__ void rec_hit(const SceneView& s, uint32_t flags, float t, uint32_t leaf, uint32_t brick,
                                        uint32_t voxel)
{
    *lane_word(s, s.aux_base, 0) = flags;
    *lane_word(s, s.aux_base, 1) = f2u(t);
    *lane_word(s, s.aux_base, 2) = leaf;
    *lane_word(s, s.aux_base, 3) = brick;
    *lane_word(s, s.aux_base, 4) = voxel;
}
// The record of the ray this lane finished, with the hit voxel's colour decoded.
__device__ __forceinline__ void rec_load(const SceneView& s, vx_hit& h)
{
    const uint32_t f = *lane_word(s, s.aux_base, 0);
    if (!(f & 1u)) {
        store_miss(h, f >> 16);
        return;
    }
    h.flags = f;
    h.t = u2f(*lane_word(s, s.aux_base, 1));
    h.leaf_node = *lane_word(s, s.aux_base, 2);
    h.brick = *lane_word(s, s.aux_base, 3);
    h.voxel = *lane_word(s, s.aux_base, 4);
    if (f & VX_HIT_FLAG_LOD) h.rgb[0] = h.rgb[1] = h.rgb[2] = 0.0f;
    else if (s.bc1) bc1_decode(ld32(s.bc1, h.brick * 32u + (h.voxel >> 4)), h.voxel & 15u, h.rgb);
    else if (s.palette) palette_decode(s.palette, ld_global<uint32_t>(reinterpret_cast<const char*>(s.brick_data + (size_t)h.brick * 512u + h.voxel)), h.rgb);
    else rgba8_decode(ld_global<uint32_t>(reinterpret_cast<const char*>(s.brick_data + (size_t)h.brick * 512u + h.voxel)), h.rgb);
}

// Per-lane LDS words after the stack: the brick-DDA inputs ol[3], dl[3],
// rinv[3], tstart, texit, then (kernels that postpone leaf visits only) a
// pending leaf's brick id and leaf descriptor index (its brick cell is
// recomputed from the unchanged ray state).
constexpr uint32_t kAuxWords = 11;
constexpr uint32_t kParkWords = 2;
// Words per lane column: stack (2 words per entry) and brick-DDA inputs; odd
// for every depth, with or without the parked-leaf words.
constexpr uint32_t lane_words(uint32_t stack_depth, bool parks)
{
    return 2u * stack_depth + kAuxWords + (parks ? kParkWords : 0u);
}
static_assert(kAuxWords % 2u == 1u && kParkWords % 2u == 0u, "lane columns must be an odd number of words apart");
constexpr uint32_t lane_lds_bytes(uint32_t stack_depth, bool parks) { return lane_words(stack_depth, parks) * 4u; }
// Lays out the per-lane columns after `staged` descriptors for a launch of
// `lanes` lanes; parks: the kernel postpones leaf visits (trace_ray with
// POSTPONE). Returns the kernel's LDS bytes.
inline uint32_t lay_out_lanes(SceneView& v, uint32_t staged, uint32_t lanes, bool parks)
{
    v.stack_base = staged * 2u;
    v.aux_base = 2u * v.stack_depth;
    v.col_words = lane_words(v.stack_depth, parks);
    return staged * 8u + lanes * lane_lds_bytes(v.stack_depth, parks);
}

// Per-lane traversal state between iterations of the ESVO loop. Holding it in
// a struct lets a kernel suspend a ray after any iteration (persistent kernels
// refill retired lanes between iterations).
struct RayState {
    float tc[3], tb[3];      // ESVO coefficients
    float pos[3];
    float t_min, t_max;
    uint32_t parent;
    int idx, scale, iter;
    int mask;                // octant mask (bits 0-2) | can-step flags (bits 3-5)
    // scale_exp2 = 2^(scale - 23), derived from scale in one instruction rather
    // than carried: one loop-carried register less to copy between branches
    __device__ __forceinline__ float scale_exp2() const { return u2f(((uint32_t)scale + 104u) << 23); }
};

// Ray set-up, coefficients and initTraversalState. Returns true when the ray
// is already finished (its miss record written); false when the loop must run.
__device__ __forceinline__ bool init_ray(const SceneView& s, const float o[3], const float d[3], float tmin,
                                         float tmax, RayState& r)
{
    rec_miss(s, 0u); // overwritten by the ray inputs when the set-up succeeds
    float ol[3], dl[3], du[3], invd[3], tstart, texit;
    if (!setup_ray(s, o, d, tmin, tmax, ol, dl, du, tstart, texit, invd)) return true;
    // brick-DDA inputs to this lane's LDS column (read back at brick visits)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        *lane_word(s, s.aux_base, i) = f2u(ol[i]);
        *lane_word(s, s.aux_base, 3 + i) = f2u(dl[i]);
        *lane_word(s, s.aux_base, 6 + i) = f2u(__builtin_fabsf(dl[i]) < 1e-8f ? (dl[i] >= 0.0f ? 1e8f : -1e8f) : invd[i]);
    }
    *lane_word(s, s.aux_base, 9) = f2u(tstart);
    *lane_word(s, s.aux_base, 10) = f2u(texit);

    float n0[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) n0[i] = (ol[i] + dl[i] * tstart) * s.invN + 1.0f;

#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float ds = du[i];
        if (__builtin_fabsf(ds) < DIR_EPS) ds = __builtin_copysignf(DIR_EPS, ds);
        r.tc[i] = 1.0f / -__builtin_fabsf(ds);
        r.tb[i] = r.tc[i] * n0[i];
    }
    int mask = 7;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (du[i] > 0.0f) {
            mask ^= 1 << i;
            r.tb[i] = 3.0f * r.tc[i] - r.tb[i];
        }

    float t_min = mx(mx(2.0f * r.tc[0] - r.tb[0], 2.0f * r.tc[1] - r.tb[1]), 2.0f * r.tc[2] - r.tb[2]);
    const float t_max = mn(mn(r.tc[0] - r.tb[0], r.tc[1] - r.tb[1]), r.tc[2] - r.tb[2]);
    t_min = mx(t_min, 0.0f);
    if (t_min >= t_max) { // the ray inputs are in the record words now: write the miss over them
        rec_miss(s, 0u);
        return true;
    }

    r.parent = 0u;
    r.scale = ESVO_MAX_SCALE;
    r.iter = 0;
    int idx = 0;
    const bool use_pos = t_min < 1e-4f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        r.pos[i] = 1.0f;
        const float m = (mask & (1 << i)) ? n0[i] : (3.0f - n0[i]);
        if (__builtin_fabsf(du[i]) < DIR_EPS || use_pos) {
            if (m >= 1.5f) {
                idx |= 1 << i;
                r.pos[i] = 1.5f;
            }
        } else if (1.5f * r.tc[i] - r.tb[i] > t_min) {
            idx ^= 1 << i;
            r.pos[i] = 1.5f;
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
        if (__builtin_fabsf(du[i]) >= DIR_EPS) mask |= 8 << i;
    r.idx = idx;
    r.mask = mask;
    r.t_min = t_min;
    r.t_max = t_max;
    return false;
}

__device__ __forceinline__ float corrected_tc_max_f(float tx, float ty, float tz, int mask, float t_max)
{
    const float thr = 1000.0f;
    const float vx = ((mask & 8) && __builtin_fabsf(tx) < thr) ? tx : t_max;
    const float vy = ((mask & 16) && __builtin_fabsf(ty) < thr) ? ty : t_max;
    const float vz = ((mask & 32) && __builtin_fabsf(tz) < thr) ? tz : t_max;
    // tx = pos*tc - tb with pos*tc <= -1 is finite and never -0; t_max is built
    // from such differences (initTraversalState) and their minima
    return fmn(fmn(vx, vy), vz);
}

// One iteration of the ESVO loop (VoxelRayMarch_Compressed.comp:396-478) is
// split at the brick visit so a wave can postpone it (trace_ray):
//   esvo_step          the iteration up to the leaf test: PUSH, or ADVANCE/POP
//                      for a skipped child, or STEP_LEAF when the child is a
//                      brick (its brick id, cell and leaf index are returned;
//                      the ray state is untouched);
//   visit_leaf         the brick DDA of a leaf (hit: the ray is done);
//   resume_after_leaf  the rest of that iteration after a brick miss (ADVANCE).
// Chained without a pause (step_ray) they are exactly the reference iteration.
enum : int { STEP_CONTINUE = 0, STEP_DONE = 1, STEP_LEAF = 2 };

// A leaf child about to be visited: brick id and the leaf descriptor's index.
struct LeafRef {
    uint32_t brick, lidx;
};

// Grid cell of the brick under the current child (SVOBrickDDA.cpp:85-93): the
// child's unmirrored corner in the unit cube times the bricks per axis.
__device__ __forceinline__ void brick_cell(const SceneView& s, const RayState& r, int bc[3])
{
    const int octm = r.mask & 7;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float lp = (octm & (1 << i)) ? r.pos[i] : (3.0f - r.scale_exp2() - r.pos[i]);
        bc[i] = (int)((lp - 1.0f) * s.Bf);
    }
}

// ADVANCE and POP of the iteration that started at r.iter, given the current
// child's exit times tx, ty, tz and corrected max tcm. Returns true when the
// ray is finished (its miss record written).
__device__ __forceinline__ bool advance_ray(const SceneView& s, RayState& r, float tx, float ty, float tz, float tcm)
{
    const int iter = r.iter;
    const float sc = r.scale_exp2();
    const bool canx = r.mask & 8, cany = r.mask & 16, canz = r.mask & 32;
    const float tca =
        tcm >= 1e10f ? fmx(fmx(canx ? tx : -1e10f, cany ? ty : -1e10f), canz ? tz : -1e10f) : tcm;
    int step_mask = 0;
    if (canx && tx <= tca) { step_mask ^= 1; r.pos[0] -= sc; }
    if (cany && ty <= tca) { step_mask ^= 2; r.pos[1] -= sc; }
    if (canz && tz <= tca) { step_mask ^= 4; r.pos[2] -= sc; }
    r.t_min = fmx(tca, 0.0f); // tca: finite, never -0 (see corrected_tc_max_f)
    r.idx ^= step_mask;
    r.iter = iter + 1;

    if ((r.idx & step_mask) == 0) {
        if (r.scale < ESVO_MAX_SCALE) r.t_max = stack_tmax(s, ESVO_MAX_SCALE - (r.scale + 1));
        return false;
    }

    // POP
    if (r.scale >= ESVO_MAX_SCALE) {
        if (r.t_min > r.t_max || r.pos[0] < 1.0f || r.pos[0] >= 2.0f || r.pos[1] < 1.0f || r.pos[1] >= 2.0f ||
            r.pos[2] < 1.0f || r.pos[2] >= 2.0f) {
            rec_miss(s, (uint32_t)(iter + 1));
            return true;
        }
        return false;
    }
    uint32_t diff = 0u;
    if (step_mask & 1) diff |= f2u(r.pos[0]) ^ f2u(r.pos[0] + sc);
    if (step_mask & 2) diff |= f2u(r.pos[1]) ^ f2u(r.pos[1] + sc);
    if (step_mask & 4) diff |= f2u(r.pos[2]) ^ f2u(r.pos[2] + sc);
    if (diff == 0u) {
        rec_miss(s, (uint32_t)(iter + 1));
        return true;
    }
    const int scale = (int)((f2u((float)diff) >> 23) - 127u);
    r.scale = scale;
    if (scale < s.min_scale || scale > ESVO_MAX_SCALE) {
        rec_miss(s, (uint32_t)(iter + 1));
        return true;
    }
    r.parent = stack_ptr(s, ESVO_MAX_SCALE - scale);
    r.t_max = stack_tmax(s, ESVO_MAX_SCALE - scale);
    const uint32_t shx = f2u(r.pos[0]) >> scale, shy = f2u(r.pos[1]) >> scale, shz = f2u(r.pos[2]) >> scale;
    r.pos[0] = u2f(shx << scale);
    r.pos[1] = u2f(shy << scale);
    r.pos[2] = u2f(shz << scale);
    r.idx = (int)((shx & 1u) | ((shy & 1u) << 1) | ((shz & 1u) << 2));
    return false;
}

// Brick DDA of a leaf child. True on a hit (its record written).
template <bool COUNT>
__device__ __forceinline__ bool visit_leaf(const SceneView& s, const RayState& r, uint32_t brick, const int bc[3],
                                           uint32_t lidx, TraceCounts& cnt)
{
    float ol[3], dl[3], rinv[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ol[i] = u2f(*lane_word(s, s.aux_base, i));
        dl[i] = u2f(*lane_word(s, s.aux_base, 3 + i));
        rinv[i] = u2f(*lane_word(s, s.aux_base, 6 + i));
    }
    const float tstart = u2f(*lane_word(s, s.aux_base, 9)), texit = u2f(*lane_word(s, s.aux_base, 10));
    uint32_t voxel, step_axis;
    if (!brick_dda<COUNT>(s, brick, bc, ol, dl, rinv, tstart, texit, voxel, step_axis, cnt)) return false;
    // reload the ray inputs for the record (the barrier keeps the compiler from
    // holding their per-axis tests live across the walk)
    asm volatile("" ::: "memory");
    float hol[3], hdl[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        hol[i] = u2f(*lane_word(s, s.aux_base, i));
        hdl[i] = u2f(*lane_word(s, s.aux_base, 3 + i));
    }
    uint32_t flags;
    float t;
    hit_entry(bc, voxel, step_axis, hol, hdl, tstart, flags, t);
    rec_hit(s, flags | ((uint32_t)(r.iter + 1) << 16), t, lidx, brick, voxel);
    if (COUNT) cnt.hits += 1;
    return true;
}

// IMMEDIATE: visit a leaf's brick inside the iteration (a miss falls through
// to ADVANCE, as in the reference); otherwise return STEP_LEAF with lf set.
// LOD: screen-space LOD termination before the leaf/PUSH decision
// (libraries/SVO/src/SVOTraversal.cpp:614-650, LODParameters::shouldTerminate
// libraries/SVO/include/SVOLOD.h:129-146,189-195): the child is accepted when the ray
// cone at distance tstart + t_min * |grid extent| is at least its world size.
template <bool COUNT, bool STAGED, bool IMMEDIATE, bool LOD = false>
__device__ __forceinline__ int esvo_step(const SceneView& s, RayState& r, TraceCounts& cnt, LeafRef& lf)
{
    const int iter = r.iter;
    // The reference loop also ends when scale exceeds the root's; here POP
    // finishes such a ray itself, so the scale is always <= ESVO_MAX_SCALE here.
    if (iter >= MAX_ITERS) {
        rec_miss(s, (uint32_t)iter);
        return STEP_DONE;
    }
    if (COUNT) cnt.iters += 1;
    // Every scene is validated at creation (vx_scene_create walks the tree from
    // the root): each child block lies inside the descriptor array and each brick
    // id is either 0xFFFFFF or a valid brick, so the parent, the leaf descriptor
    // index and the brick id below need no range checks.
    const uint2 desc = fetch_node<STAGED>(s, r.parent);
    const uint32_t valid = (desc.x >> 16) & 0xFFu, leaf = (desc.x >> 24) & 0xFFu;
    const uint32_t cp = (desc.x & 0x8000u) ? (desc.y & 0xFFFFFFu) : (desc.x & 0x7FFFu);
    const int octm = r.mask & 7;
    const int local = r.idx ^ ((~octm) & 7);

    // The child's exit times serve both the PUSH test and ADVANCE (neither pos,
    // mask nor t_max changes in between), so they are evaluated once, unconditionally.
    const float tx = r.pos[0] * r.tc[0] - r.tb[0];
    const float ty = r.pos[1] * r.tc[1] - r.tb[1];
    const float tz = r.pos[2] * r.tc[2] - r.tb[2];
    const float tcm = corrected_tc_max_f(tx, ty, tz, r.mask, r.t_max);
    const float tv_max = fmn(r.t_max, tcm); // finite, never -0 (see corrected_tc_max_f)
    const float sc = r.scale_exp2();
    const float half = sc * 0.5f;
    const float cx = half * r.tc[0] + tx;
    const float cy = half * r.tc[1] + ty;
    const float cz = half * r.tc[2] + tz;
    const bool process = ((valid >> local) & 1u) && !(r.t_min > r.t_max + T_EPS) && r.t_min <= tv_max + T_EPS;
    if (LOD && process) {
        const float dist = u2f(*lane_word(s, s.aux_base, 9)) + r.t_min * s.lod_ray_length;
        // the child's edge is scale_exp2 = 2^(scale - 23) of the unit cube (esvoScaleToWorldSize)
        if (dist * s.lod_dir + s.lod_orig >= sc * s.lod_world_size) {
            // a coarse hit at this child: the reference's placeholder normal (0,1,0),
            // its user level (esvoToUserScale) in bits 8-12, its parent descriptor
            rec_hit(s,
                    1u | (1u << 1) | (1u << 3) | VX_HIT_FLAG_LOD | ((uint32_t)(r.scale - s.min_scale) << 8) |
                        ((uint32_t)(iter + 1) << 16),
                    dist, r.parent, 0xFFFFFFFFu, 0xFFFFFFFFu);
            return STEP_DONE;
        }
    }
    if (process) {
        if ((leaf >> local) & 1u) {
            if (COUNT) cnt.leaves += 1;
            const uint32_t below = (1u << local) - 1u;
            const uint32_t lidx = cp + popc8(valid & ~leaf) + popc8(valid & leaf & below);
            const uint32_t brick = fetch_node<STAGED>(s, lidx).y & 0xFFFFFFu;
            if (brick != 0xFFFFFFu) {
                if (IMMEDIATE) {
                    int bc[3];
                    brick_cell(s, r, bc);
                    if (visit_leaf<COUNT>(s, r, brick, bc, lidx, cnt)) return STEP_DONE;
                } else {
                    lf = LeafRef{brick, lidx};
                    return STEP_LEAF;
                }
            }
            // no hit in the leaf (or no usable brick): the reference goes on to
            // ADVANCE (its t_min = tv_max is overwritten there unread)
        } else {
            stack_put(s, ESVO_MAX_SCALE - r.scale, r.parent, r.t_max);
            const uint32_t below = (1u << local) - 1u;
            r.parent = cp + popc8(valid & ~leaf & below);
            int idx = 0;
            r.scale -= 1;
            if (cx > r.t_min) { idx ^= 1; r.pos[0] += half; }
            if (cy > r.t_min) { idx ^= 2; r.pos[1] += half; }
            if (cz > r.t_min) { idx ^= 4; r.pos[2] += half; }
            r.idx = idx;
            r.t_max = tv_max;
            r.iter = iter + 1;
            return STEP_CONTINUE;
        }
    }
    return advance_ray(s, r, tx, ty, tz, tcm) ? STEP_DONE : STEP_CONTINUE;
}

// The rest of the iteration after a brick miss: ADVANCE from the leaf child
// (its exit times are recomputed from the unchanged state).
__device__ __forceinline__ bool resume_after_leaf(const SceneView& s, RayState& r)
{
    const float tx = r.pos[0] * r.tc[0] - r.tb[0];
    const float ty = r.pos[1] * r.tc[1] - r.tb[1];
    const float tz = r.pos[2] * r.tc[2] - r.tb[2];
    return advance_ray(s, r, tx, ty, tz, corrected_tc_max_f(tx, ty, tz, r.mask, r.t_max));
}

// One whole reference iteration. Returns true when the ray is finished; rec_load
// then gives its record.
template <bool COUNT, bool STAGED, bool LOD = false>
__device__ __forceinline__ bool step_ray(const SceneView& s, RayState& r, TraceCounts& cnt)
{
    LeafRef lf;
    return esvo_step<COUNT, STAGED, true, LOD>(s, r, cnt, lf) == STEP_DONE;
}

// Full traversal of one ray per lane. Writes the hit record; returns true on a
// hit. Each lane executes exactly the reference's sequence of iterations.
//
// POSTPONE (for incoherent rays): a lane whose iteration reached a brick waits
// until at least 3/4 of the wave's live lanes have, then those lanes run their
// brick DDAs together, so the wave pays for the DDA loop once per batch of
// leaves instead of once per lane's timing. Coherent camera rays reach leaves
// together anyway; for them the immediate form (no bookkeeping) is faster.
template <bool COUNT, bool STAGED, bool POSTPONE, bool LOD = false>
__device__ __forceinline__ bool trace_ray(const SceneView& s, const float o[3], const float d[3], float tmin,
                                          float tmax, vx_hit& h, TraceCounts& cnt)
{
    RayState r;
    if (!POSTPONE) {
        if (!init_ray(s, o, d, tmin, tmax, r))
            while (!step_ray<COUNT, STAGED, LOD>(s, r, cnt)) {
            }
        rec_load(s, h);
        return (h.flags & 1u) != 0u;
    }
    enum : int { ESVO = 0, LEAF = 1, DONE = 2 };
    int mode = init_ray(s, o, d, tmin, tmax, r) ? DONE : ESVO;
    for (;;) {
        if (mode == ESVO) {
            LeafRef lf;
            const int st = esvo_step<COUNT, STAGED, false, LOD>(s, r, cnt, lf);
            mode = st == STEP_DONE ? DONE : (st == STEP_LEAF ? LEAF : ESVO);
            if (mode == LEAF) { // parked in the lane's LDS column while it waits
                *lane_word(s, s.aux_base, 11) = lf.brick;
                *lane_word(s, s.aux_base, 12) = lf.lidx;
            }
        }
        const uint64_t live = __ballot(mode != DONE);
        if (live == 0ull) break;
        const uint32_t waiting = (uint32_t)__popcll(__ballot(mode == LEAF));
        // finished lanes, still in the loop, and lanes still traversing stay out
        if (waiting * 4u >= (uint32_t)__popcll(live) * 3u && mode == LEAF) {
            int bc[3];
            brick_cell(s, r, bc); // the ray state is unchanged while the leaf waits
            if (visit_leaf<COUNT>(s, r, *lane_word(s, s.aux_base, 11), bc, *lane_word(s, s.aux_base, 12), cnt))
                mode = DONE;
            else mode = resume_after_leaf(s, r) ? DONE : ESVO;
        }
    }
    rec_load(s, h);
    return (h.flags & 1u) != 0u;
}

} // namespace vx
