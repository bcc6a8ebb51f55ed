/*
 * vixen_build.h -- C ABI of the host-side scene builder and the synthetic
 * scene generators.
 *
 * The builder produces exactly the buffers vx_scene_create consumes, in VIXEN's
 * layout and order:
 *   - brick ids are Morton ranks of the populated 8^3 bricks
 *     (libraries/SVO/src/SVORebuild.cpp:334-371, Morton interleave
 *      libraries/Core/src/MortonEncoding.cpp:18-63: x in bit 0, y bit 1, z bit 2);
 *   - descriptors are in breadth-first order from the root, each node's
 *     children contiguous in octant order, leaf (brick) descriptors last
 *     (SVORebuild.cpp:419-598);
 *   - a leaf descriptor has validMask = leafMask = 0xFF and word1 = brick id
 *     (SVORebuild.cpp:397-405);
 *   - brickData is 512 u32 per brick, linear index x + 8y + 64z
 *     (libraries/CashSystem/src/VoxelSceneCacher.cpp:551-588);
 *   - BC1 blocks cover 16 consecutive linear voxels, empty voxels encoded as
 *     black, all-empty blocks stored as 0 (SVORebuild.cpp:613-731, encoder
 *     libraries/VoxelData/src/Compression/DXT1Compressor.cpp:38-111);
 *   - the 256-byte OctreeConfig as VoxelSceneCacher.cpp:592-625 fills it: grid
 *     bounds [0,N]^3 (the voxel units vx_scene_create traces in), localToWorld
 *     = scale(10) and worldToLocal its inverse scale(0.1), the compute shader's
 *     10-unit world (WORLD_GRID_SIZE, :605-625).
 * Far pointers: a child block index above 0x7FFF sets farBit and is stored in
 * word1 bits 0-23 (the contour field, unused on this path).
 */
#ifndef VIXEN_BUILD_H
#define VIXEN_BUILD_H

#include "vixen_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vx_built_scene {
    uint32_t* esvo_nodes;   /* node_count x 2 */
    uint64_t node_count;
    uint32_t* brick_data;   /* brick_count x 512 */
    uint64_t brick_count;
    uint64_t* bc1_colors;   /* brick_count x 32, NULL when not requested */
    uint32_t* brick_coords; /* brick_count: bx | by << 10 | bz << 20 (grid cell of each brick) */
    uint64_t internal_count; /* descriptors above the leaf level */
    vx_octree_config config;
} vx_built_scene;

/* What the BC1 pool is encoded from (with_bc1 of the builders):
 *   VX_BC1_NONE          no BC1 pool;
 *   VX_BC1_FROM_RGBA8    each voxel word is an RGBA8 colour (R in bits 0-7);
 *   VX_BC1_FROM_MATERIAL each voxel word is a material id, coloured by
 *                        MaterialIdToColor (vx_material_id_to_color), as
 *                        LaineKarrasOctree::rebuild encodes VIXEN's scenes
 *                        (libraries/SVO/src/SVORebuild.cpp:680-715).
 * brickData holds the words as given either way (VIXEN stores material ids,
 * VoxelSceneCacher.cpp:551-588). */
#define VX_BC1_NONE 0
#define VX_BC1_FROM_RGBA8 1
#define VX_BC1_FROM_MATERIAL 2

/* Build from a dense grid grid[x + N*(y + N*z)] of voxel words (0 = empty):
 * RGBA8 colours or material ids (see VX_BC1_*). N must be a power of two in
 * [16, 4096]. nthreads <= 0 uses all cores. */
vx_status vx_build_from_dense(const uint32_t* grid, uint32_t N, int with_bc1, int nthreads,
                              vx_built_scene** out);

/* Sparse form of vx_build_from_dense, for worlds too large for a dense grid
 * (up to 4096^3): the world as n populated bricks, as LaineKarrasOctree::rebuild
 * collects them from the voxel world (libraries/SVO/src/SVORebuild.cpp:270-330:
 * solid voxels binned by brick cell, bricks ordered by Morton code).
 * Brick i has cell (brick_cells[3i], brick_cells[3i+1], brick_cells[3i+2]) in
 * [0, N/8)^3 and RGBA8 voxels brick_voxels[512 i + z*64 + y*8 + x] (0 = empty).
 * Bricks may come in any order; bricks with no solid voxel are dropped; two
 * bricks in one cell are an error. The output equals vx_build_from_dense's on
 * the dense grid holding the same voxels. */
vx_status vx_build_from_bricks(const uint32_t* brick_cells, const uint32_t* brick_voxels, uint64_t n,
                               uint32_t N, int with_bc1, int nthreads, vx_built_scene** out);
void vx_built_scene_free(vx_built_scene* s);

/* BC1 encode of 16 colours in [0,1]^3 (DXT1Compressor::encodeBlockTyped with
 * validCount 16). */
uint64_t vx_bc1_encode_block(const float colors[16][3]);

/* MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120), the CPU table
 * the rebuild encodes BC1 colours from: getMaterialColor's constants except for
 * ids 33-39, which it lists instead of interpolating, and the same HSV wheel. */
void vx_material_id_to_color(uint32_t material_id, float rgb[3]);

/* The normal pool of VIXEN's compressed shader (compressedNormals, binding 7):
 * per brick 32 blocks of 16 consecutive linear voxels, each block two u64
 * {blockA = base normal | U selectors << 32, blockB = UV axes | V selectors << 32}
 * at normals[brick*64 + 2*block]. Per voxel the geometric normal of the brick's
 * occupancy (LaineKarrasOctree::rebuild, libraries/SVO/src/SVORebuild.cpp:130-220,
 * 613-731), encoded by DXTNormalCompressor::encodeBlockTyped
 * (libraries/VoxelData/src/Compression/DXT1Compressor.cpp:244-355); empty voxels
 * contribute (0,1,0), all-empty blocks are {0,0}. brick_data is brick_count x 512
 * words (!= 0 is occupied). nthreads <= 0 uses all cores. */
vx_status vx_encode_normals(const uint32_t* brick_data, uint64_t brick_count, int nthreads, uint64_t* normals);

/* DXTNormalCompressor::encodeBlockTyped of 16 normals (validCount 16). */
void vx_dxt_normal_encode_block(const float normals[16][3], uint64_t out[2]);

/* computeGeometricNormal (SVORebuild.cpp:173-193) of voxel (x,y,z) in one brick
 * of 512 words: -normalize(occupancy gradient), or (0,1,0) when it vanishes. */
void vx_geometric_normal(const uint32_t brick[512], int x, int y, int z, float out[3]);

/* ------------------------------------------------------------------------ */
/* VIXEN's on-disk scene cache, VSCA version 1                                */
/* (VoxelSceneCacher::SerializeToFile / DeserializeFromFile,                  */
/*  libraries/CashSystem/src/VoxelSceneCacher.cpp:179-364)                    */
/* ------------------------------------------------------------------------ */

typedef struct vx_cache vx_cache;

/* What an entry holds besides the scene buffers (VoxelSceneCreateInfo and the
 * metadata of VoxelSceneData, VoxelSceneCacher.h:145-173,222-227). materials
 * are material_count GPUMaterial records of 32 bytes, brick_grid_lookup maps
 * bx + B*by + B*B*bz to a brick id (~0u when empty); both are carried as they
 * are (the ray-cast path reads neither). Pointers stay valid until
 * vx_cache_free. */
typedef struct vx_cache_entry_info {
    uint64_t key;                /* ((sceneType*31 + resolution)*31 + (u32)(density*100))*31 + seed */
    uint32_t scene_type;         /* 0 cornell, 1 noise, 2 tunnels, 3 cityscape, 255 custom */
    uint32_t resolution;
    float density;
    uint32_t seed;
    uint32_t node_count;
    uint32_t brick_count;
    uint32_t solid_voxel_count;
    const void* materials;
    uint64_t material_count;
    const uint32_t* brick_grid_lookup;
    uint64_t brick_grid_lookup_count;
    float world_grid_size;       /* OctreeConfig::worldGridSize, the 4 bytes after the
                                    256-byte UBO (10 in every file VIXEN writes) */
} vx_cache_entry_info;

/* Reads a cache file. Errors follow the reference loader's cases: no file ->
 * VX_ERR_INVALID_ARG; bad magic, a key that does not match its create info,
 * a truncated file or buffer sizes that disagree with the node/brick counts ->
 * VX_ERR_INVALID_SCENE; another version -> VX_ERR_UNSUPPORTED (the reference
 * treats it as stale and regenerates). */
vx_status vx_cache_read(const char* path, vx_cache** out);
vx_status vx_cache_entry_count(const vx_cache* cache, uint32_t* count);
/* Entry index: its info and a vx_scene_desc pointing into the cache (ready for
 * vx_scene_create; bc1_colors / dxt_normals are NULL when the entry has none).
 * Either output may be NULL. */
vx_status vx_cache_entry(const vx_cache* cache, uint32_t index, vx_cache_entry_info* info, vx_scene_desc* desc);
vx_status vx_cache_free(vx_cache* cache);

/* One entry to write: create info, the scene buffers and the carried extras. */
typedef struct vx_cache_entry_in {
    uint32_t scene_type;
    uint32_t resolution;
    float density;
    uint32_t seed;
    const vx_scene_desc* scene;
    const void* materials;       /* material_count x 32 bytes, or NULL */
    uint64_t material_count;
    const uint32_t* brick_grid_lookup; /* or NULL */
    uint64_t brick_grid_lookup_count;
    uint32_t solid_voxel_count;
    float world_grid_size;       /* written after the config; a value that is not
                                    finite and positive (such as 0) writes VIXEN's 10 */
} vx_cache_entry_in;

/* Writes a VSCA v1 file VIXEN's VoxelSceneCacher::DeserializeFromFile loads.
 *
 * Units. A scene traces in the world box its config's grid_min/grid_max name
 * (voxel units [0,N]^3 as VIXEN's CPU castRay, VoxelSceneCacher.cpp:526-537).
 * VIXEN's compute shader instead maps world space to the unit cube with
 * worldToLocal (VoxelRayMarch_Compressed.comp:352-354), i.e. traces in
 * [0, worldGridSize]^3. To trace a cached scene with cameras and rays in that
 * frame, create it from a config copy with grid_max = world_grid_size: the
 * world -> voxel map is then N / worldGridSize, the shader's for its
 * scale-only matrices. */
vx_status vx_cache_write(const char* path, const vx_cache_entry_in* entries, uint32_t n);

/* Synthetic scenes (voxel units, world = [0,N]^3). Each fills grid (N^3 words)
 * and a camera (vx_camera, fov in degrees) for the scene's primary view.
 *   cornell: N^3 room, 1-voxel walls x=0 red, x=N-1 green, y=0, y=N-1 and
 *            z=N-1 grey 200, two grey boxes 12^3 and 12x24x12 at seeded spots,
 *            +-8 seeded noise per channel; camera at the open z=0 face centre
 *            looking +z, 60 degrees, aspect 1.
 *   cave:    solid where 4-octave fBm value noise (base frequency 1/32,
 *            lacunarity 2, gain 0.5) exceeds its (1 - occupancy) quantile;
 *            colour is a height gradient plus seeded noise; camera at a
 *            seeded empty cell, seeded look direction, 70 degrees, 16:9.
 *   urban:   (N/32)^2 lots of pitch 32, building footprints 24-30, heights
 *            uniform in [64,900] * N/1024, a lot is built with probability
 *            occupancy; 1-voxel ground; seeded colour per building; camera
 *            above the skyline looking at a seeded ground point, 70 degrees, 16:9. */
vx_status vx_gen_cornell(uint32_t N, uint64_t seed, uint32_t* grid, vx_camera* cam);
vx_status vx_gen_cave(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam);
vx_status vx_gen_urban(uint32_t N, uint64_t seed, double occupancy, uint32_t* grid, vx_camera* cam);

#ifdef __cplusplus
}
#endif

#endif
