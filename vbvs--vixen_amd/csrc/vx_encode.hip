// vx_encode.hip -- BC1 encoding of a brick pool on the device.
//
// The same encoder as the host builder's (bc1_encode16 in vx_builder.cpp, which
// follows DXT1ColorCompressor::encodeColorHead and encodeBlockTyped,
// libraries/VoxelData/src/Compression/DXT1Compressor.cpp:38-111, with
// validCount 16): per block of 16 consecutive voxels, the pair of texels
// farthest apart (first such pair in (i, j) order) gives the two RGB565 end
// points, and each texel takes the nearest of the four reference colours. FP32
// in the host code's order with no fused steps (-ffp-contract=off on both
// sides), so the blocks equal the host builder's bit for bit. The reference
// encodes on the CPU during LaineKarrasOctree::rebuild
// (libraries/SVO/src/SVORebuild.cpp:680-715); on the device a caller that edits
// a scene's brick colours in place (vx_scene_create_device) can re-encode its
// BC1 pool without a round trip to the host.
//
// One thread per block: the work is about 1500 FP32 instructions per block and
// reads 64 bytes, so the kernel is bound by VALU issue, and 32 blocks per brick
// give one wave per two bricks.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "vx_internal.hpp"

namespace vx {

constexpr uint32_t kEncodeBlock = 256;

__device__ __forceinline__ int clampi_d(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ float dist2_d(const float a[3], const float b[3])
{
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// encodeColorHead: end point c0 in bits 0-15 (B, G, R from the low bits), c1 in
// bits 16-31; each field is biased by the bits already packed below it, as the
// reference writes it.
__device__ __forceinline__ uint32_t encode_head_d(const float c0[3], const float c1[3])
{
    uint32_t head = (uint32_t)clampi_d((int)(c0[2] * 31.0f + 0.5f), 0, 31);
    head |= (uint32_t)clampi_d((int)(c0[1] * 63.0f + 0.5f - (float)head * 0.03125f), 0, 63) << 5;
    head |= (uint32_t)clampi_d((int)(c0[0] * 31.0f + 0.5f - (float)head * 0x1p-11f), 0, 31) << 11;
    head |= (uint32_t)clampi_d((int)(c1[2] * 31.0f + 0.5f - (float)head * 0x1p-16f), 0, 31) << 16;
    head |= (uint32_t)clampi_d((int)(c1[1] * 63.0f + 0.5f - (float)head * 0x1p-21f), 0, 63) << 21;
    head |= (uint32_t)clampi_d((int)(c1[0] * 31.0f + 0.5f - (float)head * 0x1p-27f), 0, 31) << 27;
    return head;
}

// The block of 16 colours; 0 when every texel is empty (word 0), as the builder
// stores it.
__device__ __forceinline__ uint64_t bc1_encode_d(const float col[16][3], bool any)
{
    if (!any) return 0ull;
    int i0 = 0, i1 = 0;
    float best = -FLT_MAX;
    for (int i = 0; i < 16; ++i)
        for (int j = i + 1; j < 16; ++j) {
            const float d = dist2_d(col[i], col[j]);
            if (d > best) {
                i0 = i;
                i1 = j;
                best = d;
            }
        }
    // the two end points, selected without indexing the colour array by a
    // lane-varying index (that would put it in scratch memory)
    float c0[3] = {0.0f, 0.0f, 0.0f}, c1[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c0[k] = i == i0 ? col[i][k] : c0[k];
            c1[k] = i == i1 ? col[i][k] : c1[k];
        }
    const uint32_t head = encode_head_d(c0, c1);
    const float S = 0x1p-32f;
    float ref[4][3];
    ref[0][0] = (float)(head << 16) * S;
    ref[0][1] = (float)(head << 21) * S;
    ref[0][2] = (float)(head << 27) * S;
    ref[1][0] = (float)head * S;
    ref[1][1] = (float)(head << 5) * S;
    ref[1][2] = (float)(head << 11) * S;
    const float c23 = 2.0f / 3.0f, c13 = 1.0f / 3.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ref[2][k] = ref[0][k] * c23 + ref[1][k] * c13;
        ref[3][k] = ref[0][k] * c13 + ref[1][k] * c23;
    }
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        uint32_t bi = 0;
        float bd = FLT_MAX;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float d = dist2_d(col[i], ref[j]);
            if (d < bd) {
                bi = (uint32_t)j;
                bd = d;
            }
        }
        bits |= bi << (2 * i);
    }
    return (uint64_t)head | ((uint64_t)bits << 32);
}

// Block g = brick * 32 + k encodes voxels 16k .. 16k + 15 of the brick (linear
// x + 8y + 64z order). A voxel word is an RGBA8 colour (R in bits 0-7, channel /
// 255 as the builder's rgba_to_float), or with a palette a material id coloured
// by palette entry id (ids below 256) or 256 + id mod 256 (larger ids); word 0
// is empty (black, and an all-empty block is 0).
__global__ void __launch_bounds__(kEncodeBlock) k_bc1_encode(const uint32_t* __restrict__ brick_data,
                                                            uint64_t n_blocks, const float* __restrict__ palette,
                                                            uint64_t* __restrict__ out)
{
    const uint64_t g = (uint64_t)blockIdx.x * kEncodeBlock + threadIdx.x;
    if (g >= n_blocks) return;
    const uint4* src = reinterpret_cast<const uint4*>(brick_data + g * 16u); // 16-byte aligned (checked)
    uint32_t w[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint4 v = src[q];
        w[4 * q + 0] = v.x;
        w[4 * q + 1] = v.y;
        w[4 * q + 2] = v.z;
        w[4 * q + 3] = v.w;
    }
    float col[16][3];
    bool any = false;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const bool solid = w[t] != 0u;
        any |= solid;
        if (!solid) {
            col[t][0] = col[t][1] = col[t][2] = 0.0f;
        } else if (palette) {
            const float* p = palette + 3u * (w[t] < 256u ? w[t] : 256u + (w[t] & 0xFFu));
            col[t][0] = p[0];
            col[t][1] = p[1];
            col[t][2] = p[2];
        } else {
            col[t][0] = (float)(w[t] & 255u) / 255.0f;
            col[t][1] = (float)((w[t] >> 8) & 255u) / 255.0f;
            col[t][2] = (float)((w[t] >> 16) & 255u) / 255.0f;
        }
    }
    out[g] = bc1_encode_d(col, any);
}

} // namespace vx

extern "C" {

vx_status vx_bc1_encode_device(const uint32_t* brick_data, uint64_t brick_count, const float* palette,
                               uint64_t* bc1_blocks, vx_stream stream)
{
    if (brick_count == 0) return VX_OK;
    if (!brick_data || !bc1_blocks) return vx::fail(VX_ERR_INVALID_ARG, "vx_bc1_encode_device: null argument");
    if ((reinterpret_cast<uintptr_t>(brick_data) & 15u) != 0u || (reinterpret_cast<uintptr_t>(bc1_blocks) & 7u) != 0u)
        return vx::fail(VX_ERR_INVALID_ARG,
                        "vx_bc1_encode_device: brick_data must be 16-byte and bc1_blocks 8-byte aligned");
    const uint64_t n_blocks = brick_count * 32u;
    const uint64_t groups = (n_blocks + vx::kEncodeBlock - 1u) / vx::kEncodeBlock;
    if (brick_count > (1ull << 40) || groups > 0x7FFFFFFFull)
        return vx::fail(VX_ERR_INVALID_ARG, "vx_bc1_encode_device: too many bricks for one launch");
    vx::k_bc1_encode<<<dim3(static_cast<unsigned>(groups)), dim3(vx::kEncodeBlock), 0,
                       static_cast<hipStream_t>(stream)>>>(brick_data, n_blocks, palette, bc1_blocks);
    if (hipGetLastError() != hipSuccess) return vx::fail(VX_ERR_HIP, "vx_bc1_encode_device: launch failed");
    return VX_OK;
}

} // extern "C"
