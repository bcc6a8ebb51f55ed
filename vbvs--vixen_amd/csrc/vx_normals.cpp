// vx_normals.cpp -- the per-brick normal pool VIXEN's compressed shader reads:
// geometric normals from voxel topology, DXT-style 128-bit normal blocks.
//
// Restated from the reference (built here with -ffp-contract=off, so the float
// arithmetic is reproducible; the reference itself builds with fast-math):
//   geometric normal   libraries/SVO/src/SVORebuild.cpp:130-220: central
//                      differences of the brick's occupancy (outside the brick
//                      counts as empty), normal = -normalize(gradient), or
//                      (0,1,0) when |gradient| <= 0.001
//   block assembly     SVORebuild.cpp:658-723: 16 consecutive linear voxels per
//                      block, empty voxels contribute (0,1,0), all 16 encoded
//                      (validCount 16), all-empty blocks stored as {0, 0}
//   block encoder      DXTNormalCompressor (libraries/VoxelData/src/Compression/
//                      DXT1Compressor.cpp:168-355): raw base normal, U axis to
//                      the farthest input, V axis to the worst U-only fit, a
//                      2-bit U and 2-bit V selector per texel
// glm::normalize(v) is v * (1 / sqrt(dot(v, v))) and glm::length(v) is
// sqrt(dot(v, v)), dot summed left to right.
#include "vixen_build.h"
#include "vx_common.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <system_error>
#include <thread>
#include <new>
#include <vector>

namespace vx {
namespace {

struct V3 {
    float x, y, z;
};
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return mul(a, 1.0f / std::sqrt(dot(a, a))); }

constexpr float kCoefs[4] = {-1.0f, -1.0f / 3.0f, 1.0f / 3.0f, 1.0f};

// encodeRawNormal (DXT1Compressor.cpp:168-195)
uint32_t encode_raw(V3 n)
{
    const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    const int axis = (ax >= std::max(ay, az)) ? 0 : (ay >= az) ? 1 : 2;
    V3 tuv;
    switch (axis) {
    case 0: tuv = n; break;
    case 1: tuv = V3{n.y, n.z, n.x}; break;
    default: tuv = V3{n.z, n.x, n.y}; break;
    }
    uint32_t r = 0;
    if (tuv.x < 0.0f) r |= 0x80000000u;
    r |= static_cast<uint32_t>(axis) << 29;
    const float t = std::fabs(tuv.x);
    if (t > 0.0f) {
        const int u = std::clamp(static_cast<int>((tuv.y / t) * 16383.0f), -0x4000, 0x3FFF);
        const int v = std::clamp(static_cast<int>((tuv.z / t) * 8191.0f), -0x2000, 0x1FFF);
        r |= static_cast<uint32_t>(u & 0x7FFF) << 14;
        r |= static_cast<uint32_t>(v & 0x3FFF);
    }
    return r;
}

// decodeRawNormal (DXT1Compressor.cpp:197-209; shaders/Compression.glsl:115-125)
V3 decode_raw(uint32_t value)
{
    const int32_t sign = static_cast<int32_t>(value) >> 31;
    const int axis = (value >> 29) & 3;
    const int32_t t = sign ^ 0x7FFFFFFF;
    const int32_t u = static_cast<int32_t>(value << 3);
    const int32_t v = static_cast<int32_t>(value << 18);
    const float ft = static_cast<float>(t), fu = static_cast<float>(u), fv = static_cast<float>(v);
    switch (axis) {
    case 0: return V3{ft, fu, fv};
    case 1: return V3{fv, ft, fu};
    default: return V3{fu, fv, ft};
    }
}

// encodeNormalAxis (DXT1Compressor.cpp:211-230)
uint32_t encode_axis(uint32_t head, V3 a, int shift)
{
    const float m = std::max({std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)});
    int e = 0;
    if (m > 0.0f) {
        uint32_t bits;
        std::memcpy(&bits, &m, 4);
        e = std::clamp(static_cast<int>(bits >> 23) + (-127 - 31 + 1 - 3 + 13), 0, 15);
    }
    const V3 s = mul(a, std::exp2(static_cast<float>(-e + (13 - 31))));
    head |= static_cast<uint32_t>(e) << shift;
    head |= static_cast<uint32_t>(
                std::clamp(static_cast<int>(s.z + 8.5f - static_cast<float>(head) * std::exp2(-4.0f - shift)), 0, 15) ^
                8)
            << (4 + shift);
    head |= static_cast<uint32_t>(
                std::clamp(static_cast<int>(s.y + 8.5f - static_cast<float>(head) * std::exp2(-8.0f - shift)), 0, 15) ^
                8)
            << (8 + shift);
    head |= static_cast<uint32_t>(
                std::clamp(static_cast<int>(s.x + 8.5f - static_cast<float>(head) * std::exp2(-12.0f - shift)), 0, 15) ^
                8)
            << (12 + shift);
    return head;
}

// decodeNormalAxis (DXT1Compressor.cpp:232-242; shaders/Compression.glsl:134-144)
V3 decode_axis(uint32_t head, int shift)
{
    const int32_t sh = static_cast<int32_t>(head << (16 - shift));
    const int e = (sh >> 16) & 15;
    const float scale = std::exp2(static_cast<float>(e + (3 - 13)));
    return V3{static_cast<float>(sh) * scale, static_cast<float>(static_cast<int32_t>(static_cast<uint32_t>(sh) << 4)) * scale,
              static_cast<float>(static_cast<int32_t>(static_cast<uint32_t>(sh) << 8)) * scale};
}

// encodeBlockTyped with validCount 16 and sequential texels (DXT1Compressor.cpp:244-355)
void encode_block(const V3 n[16], uint64_t out[2])
{
    V3 base = n[0];
    for (int i = 1; i < 16; ++i) base = add(base, n[i]);
    if (length(base) < 0.1f * 16.0f) {
        int best = 0;
        float bd = dot(n[0], base);
        for (int i = 1; i < 16; ++i) {
            const float d = dot(n[i], base);
            if (d > bd) {
                best = i;
                bd = d;
            }
        }
        base = n[best];
    }
    const uint32_t head_base = encode_raw(base);
    base = decode_raw(head_base);

    int ui = 0;
    float ud = dot(n[0], base);
    for (int i = 1; i < 16; ++i) {
        const float d = dot(n[i], base);
        if (d < ud) {
            ui = i;
            ud = d;
        }
    }
    uint32_t head_uv = encode_axis(0u, sub(mul(n[ui], length(base)), base), 0);
    const V3 u = decode_axis(head_uv, 0);

    V3 uref[4];
    for (int i = 0; i < 4; ++i) uref[i] = normalize(add(base, mul(u, kCoefs[i])));
    int vi = 0, vlerp = 0;
    float vd = FLT_MAX;
    for (int i = 0; i < 16; ++i) {
        int bl = 0;
        float bd = dot(n[i], uref[0]);
        for (int j = 1; j < 4; ++j) {
            const float d = dot(n[i], uref[j]);
            if (d > bd) {
                bl = j;
                bd = d;
            }
        }
        if (bd < vd) {
            vi = i;
            vlerp = bl;
            vd = bd;
        }
    }
    const V3 tmp = add(base, mul(u, kCoefs[vlerp]));
    head_uv = encode_axis(head_uv, sub(mul(n[vi], length(tmp)), tmp), 16);
    const V3 v = decode_axis(head_uv, 16);

    V3 uvref[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) uvref[i + j * 4] = normalize(add(add(base, mul(u, kCoefs[i])), mul(v, kCoefs[j])));
    uint32_t bits_u = 0, bits_v = 0;
    for (int i = 0; i < 16; ++i) {
        int bl = 0;
        float bd = dot(n[i], uvref[0]);
        for (int j = 1; j < 16; ++j) {
            const float d = dot(n[i], uvref[j]);
            if (d > bd) {
                bl = j;
                bd = d;
            }
        }
        bits_u |= static_cast<uint32_t>(bl & 3) << (i * 2);
        bits_v |= static_cast<uint32_t>(bl >> 2) << (i * 2);
    }
    out[0] = static_cast<uint64_t>(head_base) | (static_cast<uint64_t>(bits_u) << 32);
    out[1] = static_cast<uint64_t>(head_uv) | (static_cast<uint64_t>(bits_v) << 32);
}

// computeGeometricNormal (SVORebuild.cpp:173-193) on one brick's occupancy.
V3 geometric_normal(const uint32_t* brick, int x, int y, int z)
{
    auto occ = [brick](int i, int j, int k) -> float {
        if (i < 0 || i > 7 || j < 0 || j > 7 || k < 0 || k > 7) return 0.0f;
        return brick[i + 8 * j + 64 * k] != 0u ? 1.0f : 0.0f;
    };
    const V3 g{occ(x + 1, y, z) - occ(x - 1, y, z), occ(x, y + 1, z) - occ(x, y - 1, z),
               occ(x, y, z + 1) - occ(x, y, z - 1)};
    if (length(g) > 0.001f) {
        const V3 nn = normalize(g);
        return V3{-nn.x, -nn.y, -nn.z};
    }
    return V3{0.0f, 1.0f, 0.0f};
}

void encode_brick(const uint32_t* brick, uint64_t* out /* 64 u64 */)
{
    for (int blk = 0; blk < 32; ++blk) {
        V3 n[16];
        bool any = false;
        for (int t = 0; t < 16; ++t) {
            const int lin = blk * 16 + t;
            if (brick[lin] == 0u) {
                n[t] = V3{0.0f, 1.0f, 0.0f};
                continue;
            }
            any = true;
            n[t] = geometric_normal(brick, lin & 7, (lin >> 3) & 7, lin >> 6);
        }
        if (any) {
            encode_block(n, out + 2 * blk);
        } else {
            out[2 * blk] = 0ull;
            out[2 * blk + 1] = 0ull;
        }
    }
}

} // namespace
} // namespace vx

extern "C" {

static vx_status encode_normals_impl(const uint32_t* brick_data, uint64_t brick_count, int nthreads, uint64_t* normals)
{
    if (brick_count && (!brick_data || !normals))
        return vx::fail(VX_ERR_INVALID_ARG, "vx_encode_normals: null argument");
    if (nthreads <= 0) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>(static_cast<uint64_t>(nthreads), brick_count / 64 + 1));
    const uint64_t per = (brick_count + nt - 1) / nt;
    std::vector<std::thread> th;
    for (uint64_t t = 0; t < nt; ++t) {
        const uint64_t b = std::min(brick_count, t * per), e = std::min(brick_count, (t + 1) * per);
        auto run = [=] {
            for (uint64_t i = b; i < e; ++i) vx::encode_brick(brick_data + i * 512, normals + i * 64);
        };
        try {
            th.emplace_back(run);
        } catch (const std::system_error&) {
            run(); // no thread available: run the range here
        }
    }
    for (auto& x : th) x.join();
    return VX_OK;
}

vx_status vx_encode_normals(const uint32_t* brick_data, uint64_t brick_count, int nthreads, uint64_t* normals)
{
    try {
        return encode_normals_impl(brick_data, brick_count, nthreads, normals);
    } catch (...) {
        return vx::fail(VX_ERR_OUT_OF_MEMORY, "vx_encode_normals: out of memory");
    }
}

void vx_dxt_normal_encode_block(const float normals[16][3], uint64_t out[2])
{
    vx::V3 n[16];
    for (int i = 0; i < 16; ++i) n[i] = vx::V3{normals[i][0], normals[i][1], normals[i][2]};
    vx::encode_block(n, out);
}

void vx_geometric_normal(const uint32_t brick[512], int x, int y, int z, float out[3])
{
    const vx::V3 n = vx::geometric_normal(brick, x, y, z);
    out[0] = n.x;
    out[1] = n.y;
    out[2] = n.z;
}

} // extern "C"
