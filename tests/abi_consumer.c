/*
 * A C host of the ray-cast ABI, with no Python or PyTorch in the process: the
 * way a C or C++ caller (a render-graph node, a benchmark harness) drives the
 * library. It builds the seeded 64^3 Cornell room (seed 1, BC1 colours) with
 * the native builder, uploads it, writes the 512x512 camera rays on the
 * device, traces them with vx_trace, renders the same frame with vx_render
 * (hit records and RGBA8), and writes the trace hits, the render hits and the
 * image to the file named by argv[1], in that order. Any failure prints
 * vx_last_error() and exits with status 1.
 *
 * Built by the package build (gcc, C11) against include/ and linked with
 * libvixen_hip.so and the ROCm HIP runtime; run by tests/test_gpu_abi_consumer.py.
 */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vixen_build.h"
#include "vixen_hip.h"

#define VX_CHECK(call)                                                                         \
    do {                                                                                       \
        vx_status s_ = (call);                                                                 \
        if (s_ != VX_OK) {                                                                     \
            fprintf(stderr, "%s failed (status %d): %s\n", #call, (int)s_, vx_last_error()); \
            return 1;                                                                          \
        }                                                                                      \
    } while (0)

#define HIP_CHECK(call)                                                               \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));         \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

static int write_all(FILE* f, const void* p, size_t bytes)
{
    return fwrite(p, 1, bytes, f) == bytes ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s OUTPUT\n", argv[0]);
        return 2;
    }
    const uint32_t N = 64, W = 512, H = 512;
    const uint64_t n = (uint64_t)W * H;
    const float tmax = 1e4f;

    uint32_t* grid = (uint32_t*)malloc(sizeof(uint32_t) * N * N * N);
    if (!grid) return 1;
    vx_camera cam;
    VX_CHECK(vx_gen_cornell(N, 1, grid, &cam));
    vx_built_scene* built = NULL;
    VX_CHECK(vx_build_from_dense(grid, N, 1, 0, &built));
    free(grid);

    vx_ctx* ctx = NULL;
    VX_CHECK(vx_ctx_create(0, &ctx));
    vx_scene_desc desc;
    memset(&desc, 0, sizeof desc);
    desc.esvo_nodes = built->esvo_nodes;
    desc.node_count = built->node_count;
    desc.brick_data = built->brick_data;
    desc.brick_count = built->brick_count;
    desc.bc1_colors = built->bc1_colors;
    desc.config = &built->config;
    vx_scene* scene = NULL;
    VX_CHECK(vx_scene_create(ctx, &desc, &scene)); /* copies: the host buffers may go */
    vx_built_scene_free(built);

    hipStream_t stream;
    HIP_CHECK(hipStreamCreate(&stream));
    vx_ray* rays = NULL;
    vx_hit *trace_hits = NULL, *render_hits = NULL;
    uint32_t* image = NULL;
    HIP_CHECK(hipMalloc((void**)&rays, n * sizeof(vx_ray)));
    HIP_CHECK(hipMalloc((void**)&trace_hits, n * sizeof(vx_hit)));
    HIP_CHECK(hipMalloc((void**)&render_hits, n * sizeof(vx_hit)));
    HIP_CHECK(hipMalloc((void**)&image, n * sizeof(uint32_t)));

    VX_CHECK(vx_camera_rays(&cam, W, H, tmax, rays, (vx_stream)stream));
    VX_CHECK(vx_trace(scene, rays, n, trace_hits, (vx_stream)stream));
    VX_CHECK(vx_render(scene, &cam, W, H, tmax, render_hits, image, (vx_stream)stream));
    HIP_CHECK(hipStreamSynchronize(stream));

    vx_hit* h = (vx_hit*)malloc(n * sizeof(vx_hit));
    uint32_t* img = (uint32_t*)malloc(n * sizeof(uint32_t));
    if (!h || !img) return 1;
    FILE* f = fopen(argv[1], "wb");
    if (!f) {
        perror(argv[1]);
        return 1;
    }
    int bad = 0;
    HIP_CHECK(hipMemcpy(h, trace_hits, n * sizeof(vx_hit), hipMemcpyDeviceToHost));
    bad |= write_all(f, h, n * sizeof(vx_hit));
    HIP_CHECK(hipMemcpy(h, render_hits, n * sizeof(vx_hit), hipMemcpyDeviceToHost));
    bad |= write_all(f, h, n * sizeof(vx_hit));
    HIP_CHECK(hipMemcpy(img, image, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    bad |= write_all(f, img, n * sizeof(uint32_t));
    bad |= fclose(f) != 0;
    free(h);
    free(img);

    HIP_CHECK(hipFree(rays));
    HIP_CHECK(hipFree(trace_hits));
    HIP_CHECK(hipFree(render_hits));
    HIP_CHECK(hipFree(image));
    HIP_CHECK(hipStreamDestroy(stream));
    VX_CHECK(vx_scene_destroy(scene));
    VX_CHECK(vx_ctx_destroy(ctx));
    if (bad) {
        fprintf(stderr, "writing %s failed\n", argv[1]);
        return 1;
    }
    printf("abi_consumer: %llu rays traced and rendered\n", (unsigned long long)n);
    return 0;
}
