/*
 * vx_oracle.h -- CPU oracle for the voxel ray-cast hot path (TEST INFRASTRUCTURE).
 *
 * It checks results and is never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it. It restates, in strict IEEE FP32 C
 * (built with -ffp-contract=off, no fast-math), the reference's traversal:
 *   - ray set-up and volume entry   libraries/SVO/src/SVOTraversal.cpp:117-149,238-259,516-583
 *   - ray coefficients               libraries/SVO/src/SVOTraversal.cpp:699-728
 *   - ESVO init/PUSH/ADVANCE/POP     shaders/ESVOTraversal.glsl:84-372
 *   - main loop and t_max restore    shaders/VoxelRayMarch_Compressed.comp:396-478
 *   - leaf -> brick addressing       shaders/VoxelRayMarch_Compressed.comp:291-305
 *   - brick position from ESVO pos   libraries/SVO/src/SVOBrickDDA.cpp:85-93
 *   - brick DDA, hit t and normal    libraries/SVO/src/SVOBrickDDA.cpp:187-225,390-533
 *   - BC1 colour decode              libraries/VoxelData/src/Compression/DXT1Compressor.cpp:14-36
 *   - pinhole ray generation         shaders/RayGeneration.glsl:61-71, VoxelRayMarch_Compressed.comp:501-503
 *   - the compressed shader's shading shaders/Compression.glsl:28-41,60-75,115-177,
 *                                    VoxelRayMarch_Compressed.comp:154-165,239-256,541-548,
 *                                    shaders/Lighting.glsl:19-28
 * plus an independent FP64 dense-grid Amanatides-Woo brute force (orc_bruteforce)
 * that pins the hit voxel and t without using the octree at all.
 */
#ifndef VX_ORACLE_H
#define VX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} orc_ray;

typedef struct orc_hit {
    uint32_t flags;
    float t;
    uint32_t leaf_node;
    uint32_t brick;
    uint32_t voxel;
    float rgb[3];
} orc_hit;

typedef struct orc_scene {
    const uint32_t* nodes;      /* node_count x 2 u32 */
    uint64_t node_count;
    const uint32_t* brick_data; /* brick_count x 512 */
    uint64_t brick_count;
    const uint64_t* bc1;        /* brick_count x 32, or NULL for RGBA8 colour */
    float grid_min[3];
    float grid_max[3];
    int32_t max_levels;         /* log2(N) */
    int32_t bricks_per_axis;    /* N / 8 */
    const uint64_t* normals;    /* brick_count x 32 DXT normal blocks (2 u64 each), or NULL */
    const float* palette;       /* 256 x RGB by material id (word & 0xFF), or NULL */
} orc_scene;

/* counts[0] rays, [1] ESVO iterations, [2] brick visits, [3] DDA voxels, [4] hits */
void orc_trace(const orc_scene* s, const orc_ray* rays, uint64_t n, orc_hit* out, uint64_t counts[5]);

/* Same, over nthreads POSIX threads on contiguous ray ranges. */
void orc_trace_mt(const orc_scene* s, const orc_ray* rays, uint64_t n, orc_hit* out, uint64_t counts[5],
                  int nthreads);

/* castRayWithLOD (SVOTraversal.cpp:224-232) with LODParameters{ray_orig_size,
 * ray_dir_size}; ray_dir_size <= 0 disables LOD (then it is orc_trace). An LOD
 * stop is a hit with flags bit 6 set and the child's user level in bits 8-12. */
void orc_trace_lod(const orc_scene* s, const orc_ray* rays, uint64_t n, float ray_orig_size, float ray_dir_size,
                   orc_hit* out, int nthreads);

/* The 64-byte push constant as 16 floats/ints: pos[3], time, dir[3], fov_deg,
 * up[3], aspect, right[3], debug_mode. Writes w*h rays, row-major. */
void orc_camera_rays(const float cam[16], uint32_t w, uint32_t h, float tmax, orc_ray* out);

/* BC1 decode of one texel. */
void orc_bc1_decode(uint64_t block, int texel, float rgb[3]);

/* decodeDXT1Color as the compressed shader evaluates it (Compression.glsl:60-75). */
void orc_bc1_decode_shader(uint64_t block, int texel, float rgb[3]);

/* decodeDXTNormal (Compression.glsl:115-177), unnormalised. */
void orc_dxt_normal_decode(uint64_t block_a, uint64_t block_b, int texel, float n[3]);

/* One pixel of vx_render's RGBA8 image (R in bits 0-7, A = 255) for ray r, its
 * hit record h and the pixel's v = (py + 0.5) / height: sky on a miss; on a
 * hit computeLighting with the compressed shader's colour and normal when the
 * scene has BC1 colours and normals, else the record's colour and entry-face
 * normal. */
uint32_t orc_shade(const orc_scene* s, const orc_ray* r, const orc_hit* h, float v);

/* vx_render's RGBA8 image: camera rays, traversal and orc_shade, w*h row-major. */
void orc_render(const orc_scene* s, const float cam[16], uint32_t w, uint32_t h, float tmax, uint32_t* rgba,
                int nthreads);

/* RGBA8 colour of an uncompressed voxel word. */
void orc_rgba8_decode(uint32_t word, float rgb[3]);

/* Colour of the hit voxel (brick, linear index, its word): the BC1 texel when
 * the scene has BC1 blocks, else palette[word & 0xFF] when it has a palette,
 * else the word's RGBA8 colour. */
void orc_voxel_color(const orc_scene* s, uint32_t brick, uint32_t voxel, uint32_t word, float rgb[3]);

/* getMaterialColor (shaders/Materials.glsl:18-62) in FP32. */
void orc_material_color(uint32_t id, float rgb[3]);

/* MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120) in FP32. */
void orc_material_id_to_color(uint32_t id, float rgb[3]);

/* Ray set-up shared by the traversal and the brute force: the local-space
 * ray (voxel units, grid origin at 0), the traced segment [tstart, texit]
 * and the unit world direction. Returns 0 when the ray misses the volume or
 * is invalid. */
int orc_setup_ray(const orc_scene* s, const orc_ray* r, float ol[3], float dl[3], float du[3],
                  float* tstart, float* texit);

/* FP64 Amanatides-Woo over a dense grid (grid[x + N*(y + N*z)] != 0 is solid),
 * on the local ray orc_setup_ray produces. Per ray: vox[3] of the first solid
 * voxel (-1 on a miss), t of its AABB entry clamped to tstart, and ambiguous=1
 * when the ray passes within eps (in t) of a voxel edge or corner next to a
 * solid voxel before or at the hit (a tie that FP32 may break either way). */
void orc_bruteforce(const orc_scene* s, const uint32_t* grid, const orc_ray* rays, uint64_t n,
                    double eps, int32_t* vox, double* t, int32_t* ambiguous);

/* The reference's CPU castRay, restated verbatim for single-level worlds (at
 * most 2 bricks per axis, in voxel units): vx_cpu_reference.c. adjust = 0 is
 * the reference as written; each ORC_CPU_* bit replaces one step by the
 * oracle's choice. status[k] is 1 hit, 0 miss, -1 the ray needs a part of the
 * reference outside that scope. Returns -1 (nothing traced) for a scene
 * outside the scope. */
#define ORC_CPU_NO_UNIT_TMAX 1u
#define ORC_CPU_OCTANT_EPS 2u
#define ORC_CPU_CLIP_BRICK 4u
#define ORC_CPU_LOOKUP_INSIDE 8u
int orc_trace_cpu_reference(const orc_scene* s, const orc_ray* rays, uint64_t n, uint32_t adjust, orc_hit* out,
                            int32_t* status);

#ifdef __cplusplus
}
#endif

#endif
