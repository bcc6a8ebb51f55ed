/*
 * vixen_hip.h -- C ABI of the MI355X (gfx950) voxel ray-cast backend.
 *
 * The path behind this ABI: ESVO (Laine-Karras) sparse-voxel-octree traversal,
 * then an 8^3 brick DDA, then BC1 (DXT1) brick colour decode. Scene buffers keep
 * VIXEN's layout: 64-bit child descriptors (bit fields of
 * libraries/SVO/include/SVOTypes.h:36-117), 512 u32 per brick of `brickData`
 * (libraries/CashSystem/src/VoxelSceneCacher.cpp:551-588), 32 BC1 blocks of
 * 8 bytes per brick (libraries/SVO/src/SVORebuild.cpp:613-731) and the
 * 256-byte OctreeConfig UBO (libraries/CashSystem/include/VoxelSceneCacher.h:100-133).
 *
 * Conventions (mirroring the reference's ISVOStructure contract,
 * libraries/SVO/include/ISVOStructure.h:30-276):
 *   - every entry point returns a vx_status; nothing throws across the ABI;
 *     the text of the last failure on the calling thread is vx_last_error();
 *   - invalid rays (NaN/Inf, zero direction) produce a miss, never an error
 *     (libraries/SVO/src/SVOTraversal.cpp:238-259);
 *   - a context names one device; scenes own their device copies and may
 *     outlive the context they were created in; every call
 *     on a scene, batch or graph runs on its context's device and leaves the
 *     calling thread's current device as it found it (vx_camera_rays, which
 *     takes no scene, runs on the current device);
 *   - trace/render calls are asynchronous on the stream they are given and
 *     take DEVICE pointers for rays, hits and images (inputs already resident
 *     in HBM); scene creation takes HOST pointers and copies;
 *   - vx_trace, vx_trace_lod, vx_render and vx_camera_rays only launch kernels
 *     (no allocation, memset or synchronisation), so a caller may record them
 *     into a HIP graph with stream capture and replay the frame. Such a graph
 *     holds the scene's device pointers but no reference to the scene: keep
 *     the scene (take one with vx_scene_retain) and any buffers it borrows
 *     alive until the graph is destroyed.
 */
#ifndef VIXEN_HIP_H
#define VIXEN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hipStream_t without pulling in the HIP headers: an opaque stream handle.
 * NULL is the device's default (null) stream. */
typedef void* vx_stream;

typedef enum vx_status {
    VX_OK = 0,
    VX_ERR_INVALID_ARG = 1,
    VX_ERR_INVALID_SCENE = 2,
    VX_ERR_HIP = 3,
    VX_ERR_OUT_OF_MEMORY = 4,
    VX_ERR_UNSUPPORTED = 5
} vx_status;

/* ------------------------------------------------------------------------ */
/* Buffer layouts                                                            */
/* ------------------------------------------------------------------------ */

/* The 256-byte OctreeConfig uniform block, std140, field for field as
 * libraries/CashSystem/include/VoxelSceneCacher.h:100-133 (and the shader's
 * copy at shaders/VoxelRayMarch_Compressed.comp:60-76). Matrices are
 * column-major as glm stores them. */
typedef struct vx_octree_config {
    int32_t esvo_max_scale;      /* always 22 */
    int32_t user_max_levels;     /* log2(resolution) */
    int32_t brick_depth_levels;  /* 3 for 8^3 bricks */
    int32_t brick_size;          /* 8 */
    int32_t min_esvo_scale;      /* esvo_max_scale - user_max_levels + 1 */
    int32_t brick_esvo_scale;    /* as the reference fills it (always 20) */
    int32_t bricks_per_axis;     /* resolution / 8 */
    int32_t _padding1;
    float grid_min[3];
    float _padding2;
    float grid_max[3];
    float _padding3;
    float local_to_world[16];
    float world_to_local[16];
    float _padding4[16];
} vx_octree_config;

/* One ray: the batch form of castRay(origin, direction, tMin, tMax)
 * (libraries/SVO/include/ISVOStructure.h:122-126). 32 bytes. */
typedef struct vx_ray {
    float o[3];
    float tmin;
    float d[3];   /* need not be unit length: it is normalised as glm::normalize does */
    float tmax;
} vx_ray;

/* Hit record, 32 bytes (replaces ISVOStructure::RayHit, ISVOStructure.h:101-112).
 *   flags bit 0      : hit
 *   flags bits 1-2   : entry-face axis (0 x, 1 y, 2 z), valid on a hit
 *   flags bit 3      : entry-face normal sign is positive
 *   flags bits 4-5   : axis of the brick-DDA step into the hit voxel (1 x, 2 y,
 *                      3 z); 0 when the hit voxel is the first one the DDA tests
 *                      in its brick (the compressed shader's axisMask)
 *   flags bit 6      : LOD hit (vx_trace_lod only): the traversal stopped at a
 *                      coarse child whose projected pixel covers it; the record
 *                      then has the reference's placeholder normal (0,1,0), the
 *                      child's user level in bits 8-12, t = world distance of the
 *                      child's entry, leaf_node = its parent descriptor, brick
 *                      and voxel ~0u and rgb 0 (ISVOStructure::RayHit as built
 *                      at libraries/SVO/src/SVOTraversal.cpp:637-647)
 *   flags bits 8-12  : an LOD hit's user scale, esvoToUserScale(scale) =
 *                      scale - (23 - user_max_levels): user_max_levels - 1 for
 *                      the root's children, one less per level down
 *   flags bits 16-31 : ESVO loop iterations executed (hit or miss)
 *   t                : world-space distance to the entry of the hit voxel's AABB,
 *                      clamped to the start of the traced segment; +inf on a miss
 *   leaf_node        : index of the leaf (brick) child descriptor; ~0u on a miss
 *   brick            : brick id (Morton rank); ~0u on a miss
 *   voxel            : brick-local linear index x + 8y + 64z; ~0u on a miss
 *   rgb              : decoded colour in [0,1]; 0 on a miss                     */
typedef struct vx_hit {
    uint32_t flags;
    float t;
    uint32_t leaf_node;
    uint32_t brick;
    uint32_t voxel;
    float rgb[3];
} vx_hit;

#define VX_HIT_FLAG_HIT 1u
#define VX_HIT_FLAG_LOD (1u << 6)
#define VX_HIT_LOD_LEVEL(flags) (((flags) >> 8) & 31u)
#define VX_HIT_AXIS(flags) (((flags) >> 1) & 3u)
#define VX_HIT_ITERS(flags) ((flags) >> 16)

/* The compute pipeline's 64-byte push constant
 * (shaders/VoxelRayMarch_Compressed.comp:102-111,
 *  libraries/RenderGraph/include/Data/CameraData.h:17-27). fov is the
 * vertical field of view in degrees. */
typedef struct vx_camera {
    float pos[3];
    float time;
    float dir[3];
    float fov_deg;
    float up[3];
    float aspect;
    float right[3];
    int32_t debug_mode;
} vx_camera;

/* Scene buffers, as VoxelSceneData holds them on the CPU
 * (libraries/CashSystem/include/VoxelSceneCacher.h:190-256). */
typedef struct vx_scene_desc {
    const uint32_t* esvo_nodes;   /* node_count x {word0, word1} */
    uint64_t node_count;
    const uint32_t* brick_data;   /* brick_count x 512 u32; != 0 means occupied */
    uint64_t brick_count;
    const uint64_t* bc1_colors;   /* brick_count x 32 BC1 blocks, or NULL: colour is
                                     then brick_data's RGBA8 (R in bits 0-7) */
    const vx_octree_config* config;
    const uint64_t* dxt_normals;  /* brick_count x 32 DXT normal blocks of two u64
                                     (compressedNormals, see vx_encode_normals), or NULL */
    const float* material_palette; /* 256 x RGB, or NULL (a host pointer, copied at
                                     creation, also for vx_scene_create_device). With no
                                     BC1 colours it makes a hit's colour
                                     palette[brickData & 0xFF], the uncompressed
                                     pipeline's getMaterialColor(brickData & 0xFF)
                                     (shaders/VoxelRayMarch.comp:189-191) when it holds
                                     vx_material_palette's table; BC1 colours, when
                                     present, take precedence as in the compressed
                                     pipeline */
} vx_scene_desc;

/* getMaterialColor (shaders/Materials.glsl:18-62) of a material id, in FP32:
 * the table's constants, mix(a, b, t) = a (1 - t) + b t for ids 33-39, and the
 * HSV wheel (s = 0.7, v = 0.8) for every other id. Host-only arithmetic. */
void vx_material_color(uint32_t material_id, float rgb[3]);
/* vx_material_color of ids 0..255, RGB per id: the palette that makes
 * vx_scene_desc.material_palette give VIXEN's uncompressed colours. */
void vx_material_palette(float palette[256 * 3]);

typedef struct vx_ctx vx_ctx;
typedef struct vx_scene vx_scene;

/* ------------------------------------------------------------------------ */
/* Context and scenes                                                        */
/* ------------------------------------------------------------------------ */

/* Thread-local text of the last failure on this thread ("" if none). */
const char* vx_last_error(void);

/* Version string of this build (kernel ISA and options). */
const char* vx_version(void);

/* Replaces the Vulkan device/queue set-up around ComputeDispatchNode
 * (libraries/RenderGraph/src/Nodes/ComputeDispatchNode.cpp:134-317). */
vx_status vx_ctx_create(int device, vx_ctx** out);
vx_status vx_ctx_destroy(vx_ctx* ctx);

/* Uploads a scene (VoxelSceneCacher::UploadToGPU,
 * libraries/CashSystem/src/VoxelSceneCacher.cpp:719-891). Validates the tree:
 * every child pointer in range, leaf bits exactly at the brick-parent level,
 * brick ids below brick_count. Far pointers: when a descriptor's farBit
 * (bit 15) is set, its child block index is word1 bits 0-23 instead of the
 * 15-bit childPointer (the contour field is unused on the voxel path). */
vx_status vx_scene_create(vx_ctx* ctx, const vx_scene_desc* desc, vx_scene** out);

/* vx_scene_create without the upload: esvo_nodes, brick_data, bc1_colors and
 * dxt_normals in desc are DEVICE pointers on the context's device (8-byte
 * aligned descriptors and BC1 blocks, 16-byte aligned brick data and normal
 * blocks; config stays a host pointer). The scene uses them in place, so the
 * caller keeps them alive and unchanged until vx_scene_destroy; only the
 * derived occupancy mask and boxes (68 B per brick) are allocated. The
 * descriptors are copied to the host once for the tree checks. The device
 * form of the scene buffers that the cacher uploads
 * (VoxelSceneCacher::UploadToGPU, libraries/CashSystem/src/VoxelSceneCacher.cpp:719-891),
 * e.g. torch tensors. */
vx_status vx_scene_create_device(vx_ctx* ctx, const vx_scene_desc* desc, vx_scene** out);

/* The checks vx_scene_create runs before uploading, without touching a device. */
vx_status vx_scene_validate(const vx_scene_desc* desc);
/* Drops a reference to the scene; the last one frees it. */
vx_status vx_scene_destroy(vx_scene* scene);
/* Takes one more reference (no reference counterpart): a holder that must keep
 * the scene usable after its creator destroys it, such as an autograd graph
 * waiting for its backward pass, retains it and calls vx_scene_destroy when done. */
vx_status vx_scene_retain(vx_scene* scene);

/* Re-checks a scene whose buffers the caller changed in place, typically a
 * vx_scene_create_device scene (no reference counterpart: VIXEN re-uploads).
 * Copies the descriptors to the host, runs vx_scene_create's tree checks
 * against the scene's brick count, rebuilds the derived occupancy masks and
 * boxes from the current brick data and drops the differentiable renderer's
 * voxel rows (rebuilt on next use). If the checks fail, every later launch on
 * the scene returns VX_ERR_INVALID_SCENE until a revalidation succeeds.
 * Synchronous; nothing may run on the scene meanwhile. Batches keep the copies
 * they made at vx_batch_create. */
vx_status vx_scene_revalidate(vx_scene* scene);

/* BC1 pool of a brick pool on the current device: bc1_blocks[32 b + k] encodes
 * voxels 16k .. 16k+15 (linear x + 8y + 64z order) of brick b, exactly as the
 * host builder's pool (DXT1ColorCompressor::encodeBlockTyped with validCount
 * 16, libraries/VoxelData/src/Compression/DXT1Compressor.cpp:38-111, as
 * LaineKarrasOctree::rebuild applies it, libraries/SVO/src/SVORebuild.cpp:680-715).
 * brick_data (brick_count x 512 words, 16-byte aligned) and bc1_blocks
 * (brick_count x 32 u64, 8-byte aligned) are device pointers. A voxel word is an
 * RGBA8 colour when palette is NULL, else a material id coloured by the device
 * array palette of 512 RGB floats: entry id for ids below 256, entry
 * 256 + id mod 256 for larger ids (vx_material_id_to_color of ids 0-511 gives
 * the builder's VX_BC1_FROM_MATERIAL pool; 256 colours repeated twice give a
 * colour per id byte). Word 0 is empty and an all-empty block is 0. No reference counterpart on the device (the reference encodes on
 * the CPU): a caller that edits a vx_scene_create_device scene's colours in
 * place re-encodes its BC1 pool with it, then calls vx_scene_revalidate. */
vx_status vx_bc1_encode_device(const uint32_t* brick_data, uint64_t brick_count, const float* palette,
                               uint64_t* bc1_blocks, vx_stream stream);

/* Descriptor/brick counts and tree shape of an uploaded scene. */
vx_status vx_scene_info(const vx_scene* scene, uint64_t* node_count, uint64_t* brick_count,
                        int32_t* user_max_levels, int32_t* bricks_per_axis);

/* ------------------------------------------------------------------------ */
/* Tracing                                                                   */
/* ------------------------------------------------------------------------ */

/* Batch castRay (libraries/SVO/src/SVOTraversal.cpp:189-198 -> castRayImpl
 * :516-691, traversal as shaders/ESVOTraversal.glsl:84-372, leaf addressing as
 * shaders/VoxelRayMarch_Compressed.comp:283-305, brick DDA and hit semantics as
 * libraries/SVO/src/SVOBrickDDA.cpp:187-225,390-533, BC1 decode as
 * libraries/VoxelData/src/Compression/DXT1Compressor.cpp:14-36).
 * rays and hits are device pointers of n elements that must not overlap (a
 * call's second launch reads rays after its first has written hit records).
 * Until the call completes, hits may hold intermediate words. */
vx_status vx_trace(vx_scene* scene, const vx_ray* rays, uint64_t n, vx_hit* hits, vx_stream stream);

/* The compute pipeline's per-frame dispatch (VoxelRayMarch_Compressed.comp:493-594):
 * one pinhole ray per pixel (shaders/RayGeneration.glsl:61-71, pixel centre
 * (p + 0.5) / size, Vulkan y-flip), traced over [0, tmax]. Writes a w*h hit
 * record image (row-major, y down) when hits != NULL and/or an RGBA8 image
 * (R in bits 0-7, A = 255) when rgba8 != NULL. Misses get the sky gradient of
 * VoxelRayMarch_Compressed.comp:541-548. Hits are shaded with computeLighting
 * (shaders/Lighting.glsl:19-28):
 *   - when the scene has BC1 colours and DXT normals, as the compressed shader
 *     does: colour from decodeDXT1Color (shaders/Compression.glsl:60-75),
 *     normal = normalize(normalize(decodeDXTNormal) + 0.5 * DDA step normal),
 *     or the ray's dominant-axis normal when the hit voxel is the first the
 *     DDA tests in its brick (VoxelRayMarch_Compressed.comp:160-165,239-256);
 *   - otherwise with the hit record's colour and entry-face normal.
 * Float -> 8 bit: clamp to [0,1], then (uint)(x * 255 + 0.5). */
vx_status vx_render(vx_scene* scene, const vx_camera* camera, uint32_t width, uint32_t height,
                    float tmax, vx_hit* hits, uint32_t* rgba8, vx_stream stream);

/* The rays vx_render traces, written out (w*h vx_ray, device pointer; width
 * and height at most 65536), on the calling thread's current device. */
vx_status vx_camera_rays(const vx_camera* camera, uint32_t width, uint32_t height, float tmax,
                         vx_ray* rays, vx_stream stream);

/* One seeded cosine-weighted diffuse bounce per primary hit (the urban
 * workload's secondary rays; no reference counterpart): origin on the hit
 * voxel's entry face, 1e-3 voxel off the surface along its normal, direction
 * cosine-distributed about the normal from a hash of (seed, ray index). A
 * missed primary, or a hit record whose face-axis field (flags bits 1-2) reads 3,
 * gets a zero direction (a miss when traced). n device rays. */
vx_status vx_bounce_rays(const vx_scene* scene, const vx_ray* primary, const vx_hit* hits, uint64_t n,
                         uint32_t seed, vx_ray* out, vx_stream stream);

/* Screen-space LOD: a ray cone of diameter ray_orig_size at the origin that
 * grows by ray_dir_size per unit distance (LODParameters,
 * libraries/SVO/include/SVOLOD.h:50-177). LOD is enabled when ray_dir_size > 0. */
typedef struct vx_lod {
    float ray_orig_size;
    float ray_dir_size;
} vx_lod;

/* LODParameters::fromCamera / fromCameraWithNearPlane / withBias
 * (SVOLOD.h:83-90, 105-116, 169-175): dir = 2 tan(fov_y / screen_height / 2), orig = near * dir,
 * a bias scales both by 2^bias. fov_y in radians. Host-only arithmetic. */
vx_lod vx_lod_from_camera(float fov_y, int32_t screen_height);
vx_lod vx_lod_from_camera_near(float fov_y, int32_t screen_height, float near_plane);
vx_lod vx_lod_with_bias(vx_lod lod, float bias);

/* Batch castRayWithLOD (libraries/SVO/src/SVOTraversal.cpp:224-232; with
 * fromCamera parameters it is castRayScreenSpaceLOD, :212-222): vx_trace with
 * the termination test of castRayImpl :614-650 before every child's leaf/PUSH
 * decision: stop when (tstart + t_min |grid_max - grid_min|) * ray_dir_size +
 * ray_orig_size >= the child's edge 2^(scale-23) (grid_max.x - grid_min.x),
 * where tstart is the traced segment's start and t_min the child's entry in the
 * unit cube. Such stops are LOD hits (VX_HIT_FLAG_LOD); other records are
 * exactly vx_trace's. With LOD disabled (ray_dir_size <= 0 or NaN) this is
 * vx_trace. lod is a host pointer. */
vx_status vx_trace_lod(vx_scene* scene, const vx_ray* rays, uint64_t n, const vx_lod* lod, vx_hit* hits,
                       vx_stream stream);

/* Per-launch traversal counters for the algorithmic-bytes model:
 * counters[0] rays, [1] ESVO iterations (8-byte descriptor fetch each),
 * [2] brick visits (8-byte leaf descriptor fetch each), [3] DDA voxels examined
 * (4-byte occupancy read each), [4] hits (8-byte BC1 block or 4-byte RGBA8 read).
 * Runs the same traversal as vx_trace with counting enabled, without the
 * occupied-box pretest vx_trace uses to skip brick walks that cannot hit, so
 * the counts are the reference algorithm's (hit records are identical either
 * way); counters is a device pointer to 8 u64 that the call zeroes first. */
vx_status vx_trace_counters(vx_scene* scene, const vx_ray* rays, uint64_t n, vx_hit* hits,
                            uint64_t* counters, vx_stream stream);

/* Levels of the breadth-first descriptor array (0-4) that vx_trace/vx_render
 * stage in LDS per workgroup (a prefix of the array, at most 64 KiB).
 * Default 0: no staging, descriptors are read from global memory. */
vx_status vx_scene_set_lds_depth(vx_scene* scene, uint32_t depth);

/* ------------------------------------------------------------------------ */
/* Batched multi-scene tracing (no reference counterpart: VIXEN dispatches   */
/* one scene per compute pass, ComputeDispatchNode.cpp:249-317)              */
/* ------------------------------------------------------------------------ */

typedef struct vx_batch vx_batch;
typedef struct vx_graph vx_graph;

/* Launch shape of the persistent batched kernel. Zero fields take defaults. */
typedef struct vx_tuning {
    uint32_t waves_per_group; /* 1, 2, 4 or 8 wave64s per workgroup */
    uint32_t chunk;           /* rays per queue chunk (a chunk never spans two scenes) */
    uint32_t lds_depth;       /* breadth-first levels of every scene staged in LDS (0-4) */
    uint32_t refill_min;      /* idle lanes that trigger a refill of a wave (1-64) */
} vx_tuning;

/* Concatenates the scenes' pools (device-to-device copies) with per-scene
 * offsets. The scenes may be destroyed afterwards. */
vx_status vx_batch_create(vx_ctx* ctx, vx_scene* const* scenes, uint32_t n_scenes, vx_batch** out);
/* Drops the caller's reference; graphs captured from the batch hold their own
 * until vx_graph_destroy, and the last reference frees it. */
vx_status vx_batch_destroy(vx_batch* batch);

/* Traces the rays of all scenes in one persistent launch: scene s owns
 * rays[ray_offsets[s] .. ray_offsets[s+1]) and the same range of hits
 * (ray_offsets is a HOST array of n_scenes + 1 entries). Results are
 * bit-identical to vx_trace per scene. tuning NULL: the autotuned shape if
 * vx_batch_autotune ran, else a default. Calls on one batch must not overlap. */
vx_status vx_trace_batched(vx_batch* batch, const vx_ray* rays, const uint64_t* ray_offsets, vx_hit* hits,
                           const vx_tuning* tuning, vx_stream stream);

/* Times every shape (waves {1,2,4,8} x chunk {32..1024} x LDS depth {0..3},
 * refill at 32 idle lanes) on these rays, then refill thresholds 8, 16 and 24
 * for the fastest shape; keeps the fastest for later calls and reports it. */
vx_status vx_batch_autotune(vx_batch* batch, const vx_ray* rays, const uint64_t* ray_offsets, vx_hit* hits,
                            vx_tuning* best, float* best_ms, vx_stream stream);

/* Captures one batched launch (queue reset + persistent kernel) in a hipGraph
 * for per-frame replay on fixed buffers. The graph references the batch, which
 * stays alive until the graph is destroyed. */
vx_status vx_batch_graph_create(vx_batch* batch, const vx_ray* rays, const uint64_t* ray_offsets, vx_hit* hits,
                                const vx_tuning* tuning, vx_graph** out);
vx_status vx_graph_launch(vx_graph* graph, vx_stream stream);
vx_status vx_graph_destroy(vx_graph* graph);

/* ------------------------------------------------------------------------ */
/* Differentiable albedo rendering (no reference counterpart; the shading    */
/* is Lambertian as shaders/Lighting.glsl:19-28 without the ambient term)    */
/* ------------------------------------------------------------------------ */

#define VX_GRAD_ATOMIC 0u
#define VX_GRAD_DETERMINISTIC 1u

/* Occupied voxels of the scene = rows of its albedo array (float[rows][3]).
 * The row of a voxel is the number of occupied voxels before it in brick-id,
 * then linear (x + 8y + 64z) order. */
vx_status vx_scene_voxel_count(vx_scene* scene, uint64_t* n_voxels);

/* image[p] = albedo[row(hit voxel of pixel p)] * max(dot(n, light), 0), 0 on a
 * miss, for the pinhole rays of vx_render (n = entry-face normal, light a unit
 * vector). image is float[h*w][3]. When hit_row/hit_shade (u32/float[h*w]) are
 * given they receive each pixel's row (~0u on a miss) and shade. */
vx_status vx_albedo_forward(vx_scene* scene, const vx_camera* camera, uint32_t width, uint32_t height,
                            const float light[3], const float* albedo, float* image, uint32_t* hit_row,
                            float* hit_shade, vx_stream stream);

/* Adjoint: grad_albedo[row] += grad_image[p] * shade[p] over the pixels p that
 * hit the row. hit_row/hit_shade from vx_albedo_forward, or NULL to re-trace the
 * hits (a stored row at or above the scene's voxel count counts as a miss).
 * mode VX_GRAD_ATOMIC (FP32 atomics) or VX_GRAD_DETERMINISTIC (stable sort by
 * row, then each row's run summed in a fixed order that depends only on the
 * run's length: identical bits on every run; needs width*height < 2^31).
 * grad_albedo accumulates. The scene holds the
 * albedo calls' row table and sort workspace (built and grown on first use), so
 * albedo calls on one scene must not overlap; tracing and rendering only read it. */
vx_status vx_albedo_backward(vx_scene* scene, const vx_camera* camera, uint32_t width, uint32_t height,
                             const float light[3], const float* grad_image, const uint32_t* hit_row,
                             const float* hit_shade, float* grad_albedo, uint32_t mode, vx_stream stream);

/* No counterpart in the reference: vx_albedo_backward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training or
 * serving loops, harnesses that forbid hipMalloc). Same arguments, results
 * and modes; bit-identical gradients. It calls no hipMalloc/hipFree: the
 * deterministic mode's workspace (about 41 bytes per pixel: keys, pixel indices
 * and contributions with their sorted copies, and the sort's per-tile digit
 * counts) lives in a fixed 48 MiB device array that the library holds on
 * every device it runs on until the process ends, so width * height is at most
 * VX_NOALLOC_MAX_PIXELS (larger images: VX_ERR_INVALID_ARG). The scene's voxel
 * row table must exist already: vx_scene_voxel_count or a vx_albedo_forward on
 * the scene builds it (that call allocates); otherwise VX_ERR_INVALID_ARG.
 * A context must have been created on the scene's device. Every call on one
 * device shares the one workspace, so calls that overlap on a device (other
 * streams, other threads) race and must be serialised by the caller. */
#define VX_NOALLOC_MAX_PIXELS (1u << 20)
vx_status vx_albedo_backward_noalloc(vx_scene* scene, const vx_camera* camera, uint32_t width, uint32_t height,
                                     const float light[3], const float* grad_image, const uint32_t* hit_row,
                                     const float* hit_shade, float* grad_albedo, uint32_t mode, vx_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* VIXEN_HIP_H */
