/*
 * vx_oracle.c -- CPU oracle of the voxel ray-cast hot path. TEST INFRASTRUCTURE:
 * what the HIP kernels' results are checked against, never linked into the product.
 *
 * Built with strict IEEE FP32 (-O3 -ffp-contract=off, no fast-math). Every
 * expression keeps the evaluation order written here; the HIP kernel must match
 * it bit for bit. min/max follow glm/GLSL: min(a,b) = b < a ? b : a,
 * max(a,b) = a < b ? b : a (libraries/SVO/src/SVOTraversal.cpp uses glm::min/max
 * and std::min/max over initializer lists, which reduce left to right).
 *
 * Parity status of each function (how it is pinned; see tests/test_oracle.py):
 *   orc_setup_ray     pinned by closed-form slab tests and the FP64 brute force
 *   trace_one         pinned by orc_bruteforce (hit voxel, brick, t) on every
 *                     non-tie ray of the Cornell and cave scenes, by an FP64
 *                     slab test against every solid voxel of sparse 2048^3 and
 *                     4096^3 worlds (the deepest trees), and by hand-built
 *                     trees with known answers
 *   orc_bc1_decode    pinned exhaustively: all 4 selectors x random heads
 *                     against the integer RGB565 closed form
 *   orc_camera_rays   pinned by closed-form pixel-centre directions
 *   orc_bc1_decode_shader  pinned against orc_bc1_decode on random blocks x
 *                     texels (the two formulas differ only below 1/32 of a step)
 *   orc_dxt_normal_decode  pinned by the raw-normal bit layout and by
 *                     encode -> decode round trips (tests/test_shading.py)
 *   orc_shade         pinned by a float64 evaluation of every pixel of a frame
 *                     (within one 8-bit step)
 *   orc_trace_lod     pinned by closed forms on hand-built grids (stop level,
 *                     iteration count and distance at the entry face; the
 *                     grid-diagonal distance past it), monotonicity in the cone
 *                     width, and equality with orc_trace where LOD is disabled
 *                     or never fires (tests/test_lod.py)
 *   orc_bruteforce    textbook Amanatides-Woo in FP64; pinned by agreement with
 *                     an FP64 slab test against every solid voxel (no stepping)
 *                     on random rays over tiny dense grids
 *                     (tests/test_oracle.py::test_bruteforce_pinned_by_slab_test_per_voxel)
 *   orc_material_color, orc_material_id_to_color  pinned by the tables'
 *                     constants typed from shaders/Materials.glsl and
 *                     SVORebuild.cpp, the mix of ids 33-39 and the HSV wheel
 *                     against float64 closed forms (tests/test_materials.py)
 *   orc_trace_cpu_reference (vx_cpu_reference.c)  the reference's CPU castRay
 *                     restated verbatim for single-level worlds; equal to
 *                     trace_one on the reference's own scenario rays, and on
 *                     random rays equal to it once four named reference steps
 *                     are replaced, every replaced step siding with the FP64
 *                     brute force (tests/test_cpu_reference.py)
 */
#include "vx_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ESVO_MAX_SCALE 22
#define ESVO_STACK 23   /* shaders/ESVOTraversal.glsl:24 */
#define MAX_ITERS 512   /* shaders/ESVOTraversal.glsl:25 */
#define T_EPS 1e-6f     /* shaders/ESVOTraversal.glsl:26 */
#define DIR_EPS 1e-5f   /* shaders/ESVOTraversal.glsl:27; SVOTraversal.cpp:707 */
#define BRICK_N 8

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline float mn(float a, float b) { return (b < a) ? b : a; }
static inline float mx(float a, float b) { return (a < b) ? b : a; }
static inline int popc8(uint32_t v) { return __builtin_popcount(v & 0xFFu); }

/* ---------------------------------------------------------------------- */
/* BC1 decode: DXT1Compressor.cpp:14-36 (decodeColorHead) and :113-130     */
/* ---------------------------------------------------------------------- */
void orc_bc1_decode(uint64_t block, int texel, float rgb[3])
{
    const uint32_t head = (uint32_t)block;
    const uint32_t bits = (uint32_t)(block >> 32);
    const float S = 1.0f / 4294967296.0f; /* exactly 2^-32 */
    const uint32_t sel = (bits >> (2 * texel)) & 3u;
    const float r0 = (float)(head << 16) * S, g0 = (float)(head << 21) * S, b0 = (float)(head << 27) * S;
    const float r1 = (float)head * S, g1 = (float)(head << 5) * S, b1 = (float)(head << 11) * S;
    const float c23 = 2.0f / 3.0f, c13 = 1.0f / 3.0f;
    if (sel == 0) {
        rgb[0] = r0; rgb[1] = g0; rgb[2] = b0;
    } else if (sel == 1) {
        rgb[0] = r1; rgb[1] = g1; rgb[2] = b1;
    } else if (sel == 2) {
        rgb[0] = r0 * c23 + r1 * c13;
        rgb[1] = g0 * c23 + g1 * c13;
        rgb[2] = b0 * c23 + b1 * c13;
    } else {
        rgb[0] = r0 * c13 + r1 * c23;
        rgb[1] = g0 * c13 + g1 * c23;
        rgb[2] = b0 * c13 + b1 * c23;
    }
}

/* Uncompressed colour: RGBA8 in the brickData word, R in bits 0-7. */
void orc_rgba8_decode(uint32_t w, float rgb[3])
{
    rgb[0] = (float)(w & 255u) / 255.0f;
    rgb[1] = (float)((w >> 8) & 255u) / 255.0f;
    rgb[2] = (float)((w >> 16) & 255u) / 255.0f;
}

void orc_voxel_color(const orc_scene* s, uint32_t brick, uint32_t voxel, uint32_t word, float rgb[3])
{
    if (s->bc1) {
        orc_bc1_decode(s->bc1[(uint64_t)brick * 32u + (voxel >> 4)], (int)(voxel & 15u), rgb);
    } else if (s->palette) {
        const float* c = s->palette + 3u * (word & 0xFFu);
        rgb[0] = c[0];
        rgb[1] = c[1];
        rgb[2] = c[2];
    } else {
        orc_rgba8_decode(word, rgb);
    }
}

static void rgb3(float rgb[3], float r, float g, float b)
{
    rgb[0] = r;
    rgb[1] = g;
    rgb[2] = b;
}

/* The default branch both material tables end with: an HSV wheel, value 0.8,
 * saturation 0.7, hue (id mod 256) / 256. */
static void material_hsv(uint32_t id, float rgb[3])
{
    float hue = (float)(id % 256u) / 256.0f;
    float h = hue * 6.0f;
    int i = (int)h;
    float f = h - (float)i;
    float p = 0.8f * 0.3f;
    float q = 0.8f * (1.0f - 0.7f * f);
    float t = 0.8f * (1.0f - 0.7f * (1.0f - f));
    if (i == 0) rgb3(rgb, 0.8f, t, p);
    else if (i == 1) rgb3(rgb, q, 0.8f, p);
    else if (i == 2) rgb3(rgb, p, 0.8f, t);
    else if (i == 3) rgb3(rgb, p, q, 0.8f);
    else if (i == 4) rgb3(rgb, t, p, 0.8f);
    else rgb3(rgb, 0.8f, p, q);
}

/* The constants both tables share; 0 when the id has none. */
static int material_const(uint32_t id, float rgb[3])
{
    switch (id) {
    case 1: rgb3(rgb, 1.0f, 0.0f, 0.0f); return 1;
    case 2: rgb3(rgb, 0.0f, 1.0f, 0.0f); return 1;
    case 3: rgb3(rgb, 0.9f, 0.9f, 0.9f); return 1;
    case 4: rgb3(rgb, 0.85f, 0.85f, 0.85f); return 1;
    case 5: rgb3(rgb, 0.95f, 0.95f, 0.95f); return 1;
    case 6: rgb3(rgb, 0.8f, 0.8f, 0.8f); return 1;
    case 7: rgb3(rgb, 0.4f, 0.4f, 0.4f); return 1;
    case 10: rgb3(rgb, 0.8f, 0.6f, 0.2f); return 1;
    case 11: rgb3(rgb, 0.6f, 0.8f, 0.9f); return 1;
    case 20: rgb3(rgb, 1.0f, 0.98f, 0.9f); return 1;
    case 30: rgb3(rgb, 0.5f, 0.45f, 0.4f); return 1;
    case 31: rgb3(rgb, 0.6f, 0.55f, 0.5f); return 1;
    case 32: rgb3(rgb, 0.55f, 0.5f, 0.45f); return 1;
    case 40: rgb3(rgb, 0.8f, 0.6f, 0.2f); return 1;
    case 50: rgb3(rgb, 0.2f, 0.2f, 0.22f); return 1;
    case 60: rgb3(rgb, 0.6f, 0.58f, 0.55f); return 1;
    case 61: rgb3(rgb, 0.7f, 0.85f, 0.95f); return 1;
    default: return 0;
    }
}

void orc_material_color(uint32_t id, float rgb[3])
{
    if (material_const(id, rgb)) return;
    if (id >= 33u && id <= 39u) {
        /* mix(x, y, a) = x * (1 - a) + y * a (GLSL 4.60, 8.3) */
        float t = (float)(id - 33u) / 6.0f;
        rgb3(rgb, 0.52f * (1.0f - t) + 0.68f * t, 0.48f * (1.0f - t) + 0.6f * t, 0.42f * (1.0f - t) + 0.52f * t);
        return;
    }
    material_hsv(id, rgb);
}

void orc_material_id_to_color(uint32_t id, float rgb[3])
{
    if (material_const(id, rgb)) return;
    switch (id) {
    case 33: rgb3(rgb, 0.52f, 0.48f, 0.42f); return;
    case 34: rgb3(rgb, 0.48f, 0.44f, 0.38f); return;
    case 35: rgb3(rgb, 0.45f, 0.42f, 0.36f); return;
    case 36: rgb3(rgb, 0.58f, 0.52f, 0.46f); return;
    case 37: rgb3(rgb, 0.62f, 0.56f, 0.48f); return;
    case 38: rgb3(rgb, 0.65f, 0.58f, 0.5f); return;
    case 39: rgb3(rgb, 0.68f, 0.6f, 0.52f); return;
    default: material_hsv(id, rgb); return;
    }
}

/* ---------------------------------------------------------------------- */
/* Ray set-up: validateRayInput (SVOTraversal.cpp:238-259), intersectAABB   */
/* (:117-149) and the tMin/tMax clamp of castRayImpl (:532-537), carried    */
/* out in the grid's local voxel space [0,N]^3.                             */
/* ---------------------------------------------------------------------- */
int orc_setup_ray(const orc_scene* s, const orc_ray* r, float ol[3], float dl[3], float du[3],
                  float* tstart, float* texit)
{
    for (int i = 0; i < 3; ++i)
        if (!isfinite(r->o[i]) || !isfinite(r->d[i])) return 0;
    const float len2 = r->d[0] * r->d[0] + r->d[1] * r->d[1] + r->d[2] * r->d[2];
    if (!(len2 > 0.0f) || !isfinite(len2)) return 0;
    const float inv = 1.0f / sqrtf(len2); /* glm::normalize: v * inversesqrt(dot(v,v)) */
    for (int i = 0; i < 3; ++i) {
        du[i] = r->d[i] * inv;
        if (!isfinite(du[i])) return 0;
    }
    const int N = s->bricks_per_axis * BRICK_N;
    const float Nf = (float)N;
    const float k = Nf / (s->grid_max[0] - s->grid_min[0]); /* world -> voxel units */
    float invd[3];
    for (int i = 0; i < 3; ++i) {
        ol[i] = (r->o[i] - s->grid_min[i]) * k;
        dl[i] = du[i] * k;
        if (fabsf(dl[i]) < 1e-8f) {
            if (ol[i] < 0.0f || ol[i] > Nf) return 0;
            invd[i] = (dl[i] >= 0.0f) ? 1e20f : -1e20f;
        } else {
            invd[i] = 1.0f / dl[i];
        }
    }
    float tn[3], tf[3];
    for (int i = 0; i < 3; ++i) {
        const float t0 = (0.0f - ol[i]) * invd[i];
        const float t1 = (Nf - ol[i]) * invd[i];
        tn[i] = mn(t0, t1);
        tf[i] = mx(t0, t1);
    }
    float tEntry = mx(mx(tn[0], tn[1]), tn[2]);
    float tExit = mn(mn(tf[0], tf[1]), tf[2]);
    if (!(tEntry <= tExit && tExit >= 0.0f)) return 0;
    tEntry = mx(tEntry, r->tmin);
    tExit = mn(tExit, r->tmax);
    if (tEntry >= tExit || tExit < 0.0f) return 0;
    *tstart = mx(tEntry, 0.0f);
    *texit = tExit;
    return 1;
}

/* computeCorrectedTcMax: ESVOTraversal.glsl:168-181 */
static inline float corrected_tc_max(float tx, float ty, float tz, const float du[3], float t_max)
{
    const float thr = 1000.0f;
    const float vx = (fabsf(du[0]) >= DIR_EPS && fabsf(tx) < thr) ? tx : t_max;
    const float vy = (fabsf(du[1]) >= DIR_EPS && fabsf(ty) < thr) ? ty : t_max;
    const float vz = (fabsf(du[2]) >= DIR_EPS && fabsf(tz) < thr) ? tz : t_max;
    return mn(mn(vx, vy), vz);
}

static void set_miss(orc_hit* h, uint32_t iters)
{
    h->flags = iters << 16;
    h->t = INFINITY;
    h->leaf_node = 0xFFFFFFFFu;
    h->brick = 0xFFFFFFFFu;
    h->voxel = 0xFFFFFFFFu;
    h->rgb[0] = h->rgb[1] = h->rgb[2] = 0.0f;
}

/* Brick DDA over one 8^3 brick in local voxel space:
 * traverseBrickAndReturnHit (SVOBrickDDA.cpp:187-225) then traverseBrickView
 * (:390-533). Two additions, both to keep grazing rays consistent with the
 * geometry: the brick span is also clipped to the traced segment's end, and a
 * brick whose clipped span is empty yields no hit. Returns 1 on a hit. */
static int brick_dda(const orc_scene* s, uint32_t brick, const int bc[3], const float ol[3],
                     const float dl[3], float tstart, float texit, orc_hit* h, uint64_t* nvox)
{
    float bmin[3], invd[3], tn[3], tf[3];
    for (int i = 0; i < 3; ++i) {
        bmin[i] = (float)(bc[i] * BRICK_N);
        invd[i] = (fabsf(dl[i]) < 1e-8f) ? ((dl[i] >= 0.0f) ? 1e8f : -1e8f) : 1.0f / dl[i];
        const float bmax = bmin[i] + 8.0f;
        const float t0 = (bmin[i] - ol[i]) * invd[i];
        const float t1 = (bmax - ol[i]) * invd[i];
        tn[i] = mn(t0, t1);
        tf[i] = mx(t0, t1);
    }
    float btmin = mx(mx(tn[0], tn[1]), tn[2]);
    float btmax = mn(mn(tf[0], tf[1]), tf[2]);
    btmin = mx(btmin, tstart);
    btmax = mn(btmax, texit);
    if (btmin > btmax) return 0;

    int cur[3], step[3];
    float tDelta[3], tNext[3], entry[3];
    for (int i = 0; i < 3; ++i) {
        entry[i] = ol[i] + dl[i] * btmin;
        int c = (int)floorf(entry[i] - bmin[i]);
        cur[i] = c < 0 ? 0 : (c > BRICK_N - 1 ? BRICK_N - 1 : c);
    }
    for (int i = 0; i < 3; ++i) {
        if (fabsf(dl[i]) < 1e-8f) {
            step[i] = 0;
            tDelta[i] = FLT_MAX;
            tNext[i] = FLT_MAX;
        } else if (dl[i] > 0.0f) {
            step[i] = 1;
            tDelta[i] = 1.0f / fabsf(dl[i]);
            const float nb = bmin[i] + (float)(cur[i] + 1);
            tNext[i] = btmin + (nb - entry[i]) / dl[i];
        } else {
            step[i] = -1;
            tDelta[i] = 1.0f / fabsf(dl[i]);
            const float nb = bmin[i] + (float)cur[i];
            tNext[i] = btmin + (entry[i] - nb) / fabsf(dl[i]);
        }
    }
    /* axis of the step into the current voxel, 0 for the brick's entry voxel: the
     * compressed shader's axisMask (VoxelRayMarch_Compressed.comp:213-275) */
    uint32_t last = 0;
    for (int it = 0; it < 3 * BRICK_N; ++it) {
        if (cur[0] < 0 || cur[0] >= BRICK_N || cur[1] < 0 || cur[1] >= BRICK_N || cur[2] < 0 ||
            cur[2] >= BRICK_N)
            return 0;
        const uint32_t lin = (uint32_t)(cur[0] + 8 * cur[1] + 64 * cur[2]);
        ++*nvox;
        const uint32_t w = s->brick_data[(uint64_t)brick * 512u + lin];
        if (w != 0u) {
            float tnh[3];
            for (int i = 0; i < 3; ++i) {
                if (fabsf(dl[i]) < 1e-8f) {
                    tnh[i] = -INFINITY;
                } else {
                    const float vmin = bmin[i] + (float)cur[i];
                    const float vmax = vmin + 1.0f;
                    const float t0 = (vmin - ol[i]) / dl[i];
                    const float t1 = (vmax - ol[i]) / dl[i];
                    tnh[i] = mn(t0, t1);
                }
            }
            float tHit = mx(mx(tnh[0], tnh[1]), tnh[2]);
            tHit = mx(tHit, tstart);
            uint32_t axis;
            if (tnh[0] >= tnh[1] && tnh[0] >= tnh[2]) axis = 0;
            else if (tnh[1] >= tnh[2]) axis = 1;
            else axis = 2;
            const uint32_t signpos = (dl[axis] > 0.0f) ? 0u : 1u;
            h->flags = 1u | (axis << 1) | (signpos << 3) | (last << 4);
            h->t = tHit;
            h->brick = brick;
            h->voxel = lin;
            orc_voxel_color(s, brick, lin, w, h->rgb);
            return 1;
        }
        if (tNext[0] < tNext[1] && tNext[0] < tNext[2]) {
            if (tNext[0] > btmax) return 0;
            cur[0] += step[0];
            tNext[0] += tDelta[0];
            last = 1;
        } else if (tNext[1] < tNext[2]) {
            if (tNext[1] > btmax) return 0;
            cur[1] += step[1];
            tNext[1] += tDelta[1];
            last = 2;
        } else {
            if (tNext[2] > btmax) return 0;
            cur[2] += step[2];
            tNext[2] += tDelta[2];
            last = 3;
        }
    }
    return 0;
}

/* lod: NULL, or {ray_orig_size, ray_dir_size} of an enabled LODParameters */
static void trace_one(const orc_scene* s, const orc_ray* r, orc_hit* h, uint64_t c[5], const float* lod)
{
    c[0] += 1;
    set_miss(h, 0);
    float ol[3], dl[3], du[3], tstart, texit;
    if (!orc_setup_ray(s, r, ol, dl, du, &tstart, &texit)) return;

    const int N = s->bricks_per_axis * BRICK_N;
    const float invN = 1.0f / (float)N;
    const float B = (float)s->bricks_per_axis;
    const int min_scale = ESVO_MAX_SCALE - s->max_levels + 1;

    /* castRayImpl :545-548: segment start mapped into the [1,2]^3 cube */
    float n0[3];
    for (int i = 0; i < 3; ++i) n0[i] = (ol[i] + dl[i] * tstart) * invN + 1.0f;

    /* computeRayCoefficients, SVOTraversal.cpp:699-728 */
    float tc[3], tb[3];
    for (int i = 0; i < 3; ++i) {
        float ds = du[i];
        if (fabsf(ds) < DIR_EPS) ds = copysignf(DIR_EPS, ds);
        tc[i] = 1.0f / -fabsf(ds);
        tb[i] = tc[i] * n0[i];
    }
    int mask = 7;
    for (int i = 0; i < 3; ++i)
        if (du[i] > 0.0f) {
            mask ^= 1 << i;
            tb[i] = 3.0f * tc[i] - tb[i];
        }

    /* initTraversalState, ESVOTraversal.glsl:84-152 */
    float t_min = mx(mx(2.0f * tc[0] - tb[0], 2.0f * tc[1] - tb[1]), 2.0f * tc[2] - tb[2]);
    float t_max = mn(mn(tc[0] - tb[0], tc[1] - tb[1]), tc[2] - tb[2]);
    t_min = mx(t_min, 0.0f);
    if (t_min >= t_max) return; /* VoxelRayMarch_Compressed.comp:386-394 */

    uint32_t stack_ptr[ESVO_STACK];
    float stack_tmax[ESVO_STACK];
    for (int i = 0; i < ESVO_STACK; ++i) {
        stack_ptr[i] = 0u;
        stack_tmax[i] = t_max;
    }
    uint32_t parent = 0u;
    int scale = ESVO_MAX_SCALE;
    float scale_exp2 = 0.5f;
    float pos[3] = {1.0f, 1.0f, 1.0f};
    int idx = 0;
    const int use_pos = t_min < 1e-4f;
    for (int i = 0; i < 3; ++i) {
        const float m = (mask & (1 << i)) ? n0[i] : (3.0f - n0[i]);
        if (fabsf(du[i]) < DIR_EPS || use_pos) {
            if (m >= 1.5f) {
                idx |= 1 << i;
                pos[i] = 1.5f;
            }
        } else if (1.5f * tc[i] - tb[i] > t_min) {
            idx ^= 1 << i;
            pos[i] = 1.5f;
        }
    }

    /* main loop: VoxelRayMarch_Compressed.comp:396-478 */
    int iter;
    for (iter = 0; iter < MAX_ITERS && scale <= ESVO_MAX_SCALE; ++iter) {
        c[1] += 1;
        if ((uint64_t)parent >= s->node_count) {
            set_miss(h, (uint32_t)(iter + 1));
            return;
        }
        const uint32_t w0 = s->nodes[2u * parent], w1 = s->nodes[2u * parent + 1u];
        const uint32_t valid = (w0 >> 16) & 0xFFu, leaf = (w0 >> 24) & 0xFFu;
        const uint32_t cp = (w0 & 0x8000u) ? (w1 & 0xFFFFFFu) : (w0 & 0x7FFFu);
        const int local = idx ^ ((~mask) & 7);

        /* checkChildValidity, ESVOTraversal.glsl:189-219 */
        int process = 0;
        float tv_max = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        if (((valid >> local) & 1u) && !(t_min > t_max + T_EPS)) {
            const float tx = pos[0] * tc[0] - tb[0];
            const float ty = pos[1] * tc[1] - tb[1];
            const float tz = pos[2] * tc[2] - tb[2];
            const float tcm = corrected_tc_max(tx, ty, tz, du, t_max);
            tv_max = mn(t_max, tcm);
            const float half = scale_exp2 * 0.5f;
            cx = half * tc[0] + tx;
            cy = half * tc[1] + ty;
            cz = half * tc[2] + tz;
            process = t_min <= tv_max + T_EPS;
        }
        if (process && lod) {
            /* screen-space LOD termination, castRayImpl SVOTraversal.cpp:614-650:
             * worldDistance = tRayStart + t_min * length(worldSize) (:545-552),
             * worldVoxelSize = esvoScaleToWorldSize(scale, worldSize.x)
             * (SVOLOD.h:189-195), terminate when worldDistance * rayDirSize +
             * rayOrigSize >= worldVoxelSize (SVOLOD.h:129-146) */
            const float ex = s->grid_max[0] - s->grid_min[0];
            const float ey = s->grid_max[1] - s->grid_min[1];
            const float ez = s->grid_max[2] - s->grid_min[2];
            const float world_len = sqrtf((ex * ex + ey * ey) + ez * ez);
            const float dist = tstart + t_min * world_len;
            const float size = ldexpf(1.0f, scale - ESVO_MAX_SCALE - 1) * ex;
            if (dist * lod[1] + lod[0] >= size) {
                /* the LOD RayHit of :637-647: normal (0,1,0), tMin = worldDistance,
                 * scale = esvoToUserScale(scale) */
                h->flags = 1u | (1u << 1) | (1u << 3) | (1u << 6) | ((uint32_t)(scale - min_scale) << 8) |
                           ((uint32_t)(iter + 1) << 16);
                h->t = dist;
                h->leaf_node = parent;
                h->brick = 0xFFFFFFFFu;
                h->voxel = 0xFFFFFFFFu;
                h->rgb[0] = h->rgb[1] = h->rgb[2] = 0.0f;
                return;
            }
        }
        if (process) {
            if ((leaf >> local) & 1u) {
                /* leaf: VoxelRayMarch_Compressed.comp:291-305 for the descriptor,
                 * SVOBrickDDA.cpp:85-93 for the brick's grid cell */
                c[2] += 1;
                const uint32_t below = (1u << local) - 1u;
                const uint32_t lidx = cp + (uint32_t)popc8(valid & ~leaf) + (uint32_t)popc8(valid & leaf & below);
                if ((uint64_t)lidx < s->node_count) {
                    const uint32_t brick = s->nodes[2u * lidx + 1u] & 0xFFFFFFu;
                    if (brick != 0xFFFFFFu && (uint64_t)brick < s->brick_count) {
                        int bc[3];
                        for (int i = 0; i < 3; ++i) {
                            const float lp = (mask & (1 << i)) ? pos[i] : (3.0f - scale_exp2 - pos[i]);
                            bc[i] = (int)((lp - 1.0f) * B);
                        }
                        if (brick_dda(s, brick, bc, ol, dl, tstart, texit, h, &c[3])) {
                            h->flags |= (uint32_t)(iter + 1) << 16;
                            h->leaf_node = lidx;
                            c[4] += 1;
                            return;
                        }
                    }
                }
                t_min = tv_max;
            } else {
                /* PUSH, ESVOTraversal.glsl:225-263 (always pushes; h is never read) */
                if (scale >= 0 && scale < ESVO_STACK) {
                    stack_ptr[scale] = parent;
                    stack_tmax[scale] = t_max;
                }
                const uint32_t below = (1u << local) - 1u;
                parent = cp + (uint32_t)popc8(valid & ~leaf & below);
                idx = 0;
                scale -= 1;
                const float half = scale_exp2 * 0.5f;
                scale_exp2 = half;
                if (cx > t_min) { idx ^= 1; pos[0] += scale_exp2; }
                if (cy > t_min) { idx ^= 2; pos[1] += scale_exp2; }
                if (cz > t_min) { idx ^= 4; pos[2] += scale_exp2; }
                t_max = tv_max;
                continue;
            }
        }

        /* ADVANCE, ESVOTraversal.glsl:270-306 */
        const float tx = pos[0] * tc[0] - tb[0];
        const float ty = pos[1] * tc[1] - tb[1];
        const float tz = pos[2] * tc[2] - tb[2];
        const int canx = fabsf(du[0]) >= DIR_EPS, cany = fabsf(du[1]) >= DIR_EPS, canz = fabsf(du[2]) >= DIR_EPS;
        float tcm = corrected_tc_max(tx, ty, tz, du, t_max);
        if (tcm >= 1e10f) tcm = mx(mx(canx ? tx : -1e10f, cany ? ty : -1e10f), canz ? tz : -1e10f);
        int step_mask = 0;
        if (canx && tx <= tcm) { step_mask ^= 1; pos[0] -= scale_exp2; }
        if (cany && ty <= tcm) { step_mask ^= 2; pos[1] -= scale_exp2; }
        if (canz && tz <= tcm) { step_mask ^= 4; pos[2] -= scale_exp2; }
        t_min = mx(tcm, 0.0f);
        idx ^= step_mask;

        if ((idx & step_mask) == 0) {
            /* VoxelRayMarch_Compressed.comp:455-459 */
            if (scale < ESVO_MAX_SCALE) t_max = stack_tmax[scale + 1];
            continue;
        }

        /* POP, ESVOTraversal.glsl:315-372 */
        if (scale >= ESVO_MAX_SCALE) {
            if (t_min > t_max || pos[0] < 1.0f || pos[0] >= 2.0f || pos[1] < 1.0f || pos[1] >= 2.0f ||
                pos[2] < 1.0f || pos[2] >= 2.0f) {
                set_miss(h, (uint32_t)(iter + 1));
                return;
            }
            continue;
        }
        uint32_t diff = 0u;
        if (step_mask & 1) diff |= f2u(pos[0]) ^ f2u(pos[0] + scale_exp2);
        if (step_mask & 2) diff |= f2u(pos[1]) ^ f2u(pos[1] + scale_exp2);
        if (step_mask & 4) diff |= f2u(pos[2]) ^ f2u(pos[2] + scale_exp2);
        if (diff == 0u) {
            set_miss(h, (uint32_t)(iter + 1));
            return;
        }
        scale = (int)((f2u((float)diff) >> 23) - 127u);
        scale_exp2 = u2f((uint32_t)(scale - ESVO_MAX_SCALE - 1 + 127) << 23);
        if (scale < min_scale || scale > ESVO_MAX_SCALE) {
            set_miss(h, (uint32_t)(iter + 1));
            return;
        }
        parent = stack_ptr[scale];
        t_max = stack_tmax[scale];
        const uint32_t shx = f2u(pos[0]) >> scale, shy = f2u(pos[1]) >> scale, shz = f2u(pos[2]) >> scale;
        pos[0] = u2f(shx << scale);
        pos[1] = u2f(shy << scale);
        pos[2] = u2f(shz << scale);
        idx = (int)((shx & 1u) | ((shy & 1u) << 1) | ((shz & 1u) << 2));
    }
    set_miss(h, (uint32_t)iter);
}

void orc_trace(const orc_scene* s, const orc_ray* rays, uint64_t n, orc_hit* out, uint64_t counts[5])
{
    uint64_t c[5] = {0, 0, 0, 0, 0};
    for (uint64_t i = 0; i < n; ++i) trace_one(s, &rays[i], &out[i], c, NULL);
    if (counts)
        for (int k = 0; k < 5; ++k) counts[k] += c[k];
}

typedef struct {
    const orc_scene* s;
    const orc_ray* rays;
    orc_hit* out;
    uint64_t begin, end;
    uint64_t c[5];
    const float* lod;
} orc_job;

static void* orc_worker(void* p)
{
    orc_job* j = (orc_job*)p;
    memset(j->c, 0, sizeof j->c);
    for (uint64_t i = j->begin; i < j->end; ++i) trace_one(j->s, &j->rays[i], &j->out[i], j->c, j->lod);
    return NULL;
}

static void trace_mt(const orc_scene* s, const orc_ray* rays, uint64_t n, orc_hit* out, uint64_t counts[5],
                     int nthreads, const float* lod)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    orc_job jobs[256];
    pthread_t th[256];
    int started[256];
    const uint64_t per = (n + (uint64_t)nthreads - 1) / (uint64_t)nthreads;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].s = s;
        jobs[t].rays = rays;
        jobs[t].out = out;
        jobs[t].lod = lod;
        jobs[t].begin = per * (uint64_t)t < n ? per * (uint64_t)t : n;
        jobs[t].end = per * (uint64_t)(t + 1) < n ? per * (uint64_t)(t + 1) : n;
        started[t] = pthread_create(&th[t], NULL, orc_worker, &jobs[t]) == 0;
        if (!started[t]) orc_worker(&jobs[t]); /* no thread: run the range here */
    }
    for (int t = 0; t < nthreads; ++t) {
        if (started[t]) pthread_join(th[t], NULL);
        if (counts)
            for (int k = 0; k < 5; ++k) counts[k] += jobs[t].c[k];
    }
}

void orc_trace_mt(const orc_scene* s, const orc_ray* rays, uint64_t n, orc_hit* out, uint64_t counts[5],
                  int nthreads)
{
    trace_mt(s, rays, n, out, counts, nthreads, NULL);
}

void orc_trace_lod(const orc_scene* s, const orc_ray* rays, uint64_t n, float ray_orig_size, float ray_dir_size,
                   orc_hit* out, int nthreads)
{
    const float lod[2] = {ray_orig_size, ray_dir_size};
    trace_mt(s, rays, n, out, NULL, nthreads, ray_dir_size > 0.0f ? lod : NULL); /* LODParameters::isEnabled */
}

/* ---------------------------------------------------------------------- */
/* The compressed shader's per-pixel shading                               */
/* ---------------------------------------------------------------------- */

/* decodeDXT1Color (Compression.glsl:28-40, 60-75) */
void orc_bc1_decode_shader(uint64_t block, int texel, float rgb[3])
{
    const float DXT_SCALE = 1.0f / 4294967296.0f;
    const float coefs[4] = {DXT_SCALE, 0.0f, 2.0f * DXT_SCALE / 3.0f, 1.0f * DXT_SCALE / 3.0f};
    const uint32_t head = (uint32_t)block, bits = (uint32_t)(block >> 32);
    const float c0 = coefs[(bits >> (texel * 2)) & 3u];
    const float c1 = DXT_SCALE - c0;
    rgb[0] = c0 * (float)(head << 16) + c1 * (float)head;
    rgb[1] = c0 * (float)(head << 21) + c1 * (float)(head << 5);
    rgb[2] = c0 * (float)(head << 27) + c1 * (float)(head << 11);
}

/* decodeRawNormal (Compression.glsl:115-125) */
static void decode_raw_normal(uint32_t value, float n[3])
{
    const int32_t sign = (int32_t)value >> 31;
    const int axis = (int)((value >> 29) & 3u);
    const int32_t t = sign ^ 0x7FFFFFFF;
    const int32_t u = (int32_t)(value << 3);
    const int32_t v = (int32_t)(value << 18);
    if (axis == 0) { n[0] = (float)t; n[1] = (float)u; n[2] = (float)v; return; }
    if (axis == 1) { n[0] = (float)v; n[1] = (float)t; n[2] = (float)u; return; }
    n[0] = (float)u; n[1] = (float)v; n[2] = (float)t;
}

/* decodeNormalAxis (Compression.glsl:134-144) */
static void decode_normal_axis(uint32_t head_uv, int shift, float a[3])
{
    const int32_t shifted = (int32_t)(head_uv << (16 - shift));
    const int exponent = (shifted >> 16) & 15;
    const float scale = exp2f((float)(exponent + (3 - 13)));
    a[0] = (float)shifted * scale;
    a[1] = (float)(int32_t)((uint32_t)shifted << 4) * scale;
    a[2] = (float)(int32_t)((uint32_t)shifted << 8) * scale;
}

/* decodeDXTNormal (Compression.glsl:41, 159-177) */
void orc_dxt_normal_decode(uint64_t block_a, uint64_t block_b, int texel, float n[3])
{
    const float coefs[4] = {-1.0f, -1.0f / 3.0f, 1.0f / 3.0f, 1.0f};
    float base[3], u[3], v[3];
    decode_raw_normal((uint32_t)block_a, base);
    decode_normal_axis((uint32_t)block_b, 0, u);
    decode_normal_axis((uint32_t)block_b, 16, v);
    const float cu = coefs[((uint32_t)(block_a >> 32) >> (texel * 2)) & 3u];
    const float cv = coefs[((uint32_t)(block_b >> 32) >> (texel * 2)) & 3u];
    for (int i = 0; i < 3; ++i) n[i] = base[i] + u[i] * cu + v[i] * cv;
}

/* GLSL normalize as v * (1 / sqrt(dot(v, v))) */
static void normalize3(float v[3])
{
    const float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; ++i) v[i] = v[i] * inv;
}

static uint32_t to_u8(float x)
{
    x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    return (uint32_t)(x * 255.0f + 0.5f);
}

static float glsl_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

uint32_t orc_shade(const orc_scene* s, const orc_ray* r, const orc_hit* h, float v)
{
    float col[3];
    if (!(h->flags & 1u)) {
        /* sky: vec3(0.5, 0.7, 1.0) * (1.0 - uv.y * 0.5) (VoxelRayMarch_Compressed.comp:544) */
        const float g = 1.0f - v * 0.5f;
        col[0] = 0.5f * g;
        col[1] = 0.7f * g;
        col[2] = 1.0f * g;
    } else {
        /* computeLighting: lightDir = normalize(vec3(1, 1, -1)), folded to the
         * correctly rounded 1/sqrt(3); color * (0.3 + max(dot(n, l), 0) * 0.7) */
        const float L[3] = {0.57735026918962576f, 0.57735026918962576f, -0.57735026918962576f};
        float rgb[3], n[3] = {0.0f, 0.0f, 0.0f};
        if (s->normals && s->bc1) {
            float ol[3], dl[3], du[3], ts, te;
            orc_setup_ray(s, r, ol, dl, du, &ts, &te); /* the normalised world direction */
            const uint64_t blk = (uint64_t)h->brick * 32u + (h->voxel >> 4);
            orc_bc1_decode_shader(s->bc1[blk], (int)(h->voxel & 15u), rgb);
            const uint32_t axis_mask = (h->flags >> 4) & 3u; /* 0: no DDA step in the hit brick */
            if (axis_mask == 0u) {
                /* VoxelRayMarch_Compressed.comp:241-249 */
                const float ax = fabsf(du[0]), ay = fabsf(du[1]), az = fabsf(du[2]);
                if (ax > ay && ax > az) n[0] = -glsl_sign(du[0]);
                else if (ay > az) n[1] = -glsl_sign(du[1]);
                else n[2] = -glsl_sign(du[2]);
            } else {
                /* :160-165 and :250-255: normalize(normalize(decoded) + ddaNormal * 0.5) */
                orc_dxt_normal_decode(s->normals[2 * blk], s->normals[2 * blk + 1], (int)(h->voxel & 15u), n);
                normalize3(n);
                float dda[3] = {0.0f, 0.0f, 0.0f};
                const int a = (int)axis_mask - 1;
                const float st = du[a] < 0.0f ? -1.0f : 1.0f; /* ivec3(sign(rayDir)), 0 -> 1 */
                dda[a] = -st;
                for (int i = 0; i < 3; ++i) n[i] = n[i] + dda[i] * 0.5f;
                normalize3(n);
            }
        } else {
            const uint32_t axis = (h->flags >> 1) & 3u;
            n[axis] = (h->flags & 8u) ? 1.0f : -1.0f; /* entry-face normal */
            for (int i = 0; i < 3; ++i) rgb[i] = h->rgb[i];
        }
        const float d = n[0] * L[0] + n[1] * L[1] + n[2] * L[2];
        const float k = 0.3f + (d < 0.0f ? 0.0f : d) * 0.7f;
        for (int i = 0; i < 3; ++i) col[i] = rgb[i] * k;
    }
    return to_u8(col[0]) | (to_u8(col[1]) << 8) | (to_u8(col[2]) << 16) | 0xFF000000u;
}

void orc_render(const orc_scene* s, const float cam[16], uint32_t w, uint32_t h, float tmax, uint32_t* rgba,
                int nthreads)
{
    const uint64_t n = (uint64_t)w * h;
    orc_ray* rays = (orc_ray*)malloc(n * sizeof(orc_ray));
    orc_hit* hits = (orc_hit*)malloc(n * sizeof(orc_hit));
    if (!rays || !hits) {
        free(rays);
        free(hits);
        return;
    }
    orc_camera_rays(cam, w, h, tmax, rays);
    orc_trace_mt(s, rays, n, hits, NULL, nthreads);
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            const uint64_t i = (uint64_t)py * w + px;
            rgba[i] = orc_shade(s, &rays[i], &hits[i], ((float)py + 0.5f) / (float)h);
        }
    free(rays);
    free(hits);
}

/* Pinhole rays: getRayDir (RayGeneration.glsl:61-71) at pixel centres
 * uv = (p + 0.5) / size (VoxelRayMarch_Compressed.comp:501-503). */
void orc_camera_rays(const float cam[16], uint32_t w, uint32_t h, float tmax, orc_ray* out)
{
    const float* pos = cam + 0;
    const float* dir = cam + 4;
    const float fov = cam[7];
    const float* up = cam + 8;
    const float aspect = cam[11];
    const float* right = cam + 12;
    const float tanHalf = tanf((fov * 0.5f) * 0.017453292519943295f);
    for (uint32_t py = 0; py < h; ++py) {
        for (uint32_t px = 0; px < w; ++px) {
            const float u = ((float)px + 0.5f) / (float)w;
            const float v = ((float)py + 0.5f) / (float)h;
            const float ndcx = u * 2.0f - 1.0f;
            const float ndcy = -(v * 2.0f - 1.0f);
            orc_ray* r = &out[(uint64_t)py * w + px];
            for (int i = 0; i < 3; ++i) {
                const float a = ((right[i] * ndcx) * tanHalf) * aspect;
                const float b = (up[i] * ndcy) * tanHalf;
                r->o[i] = pos[i];
                r->d[i] = (dir[i] + a) + b;
            }
            r->tmin = 0.0f;
            r->tmax = tmax;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* FP64 dense-grid brute force (Amanatides & Woo 1987), independent of the  */
/* octree. It walks every voxel the local ray crosses from tstart to texit. */
/* ---------------------------------------------------------------------- */
static int solid_at(const uint32_t* grid, int N, int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= N || y >= N || z >= N) return 0;
    return grid[(uint64_t)x + (uint64_t)N * ((uint64_t)y + (uint64_t)N * (uint64_t)z)] != 0u;
}

void orc_bruteforce(const orc_scene* s, const uint32_t* grid, const orc_ray* rays, uint64_t n, double eps,
                    int32_t* vox, double* tout, int32_t* amb)
{
    const int N = s->bricks_per_axis * BRICK_N;
    for (uint64_t r = 0; r < n; ++r) {
        vox[3 * r] = vox[3 * r + 1] = vox[3 * r + 2] = -1;
        tout[r] = INFINITY;
        amb[r] = 0;
        float olf[3], dlf[3], du[3], tsf, tef;
        if (!orc_setup_ray(s, &rays[r], olf, dlf, du, &tsf, &tef)) continue;
        double o[3], d[3], tNext[3], tDelta[3];
        int cur[3], step[3];
        const double ts = tsf, te = tef;
        for (int i = 0; i < 3; ++i) {
            o[i] = olf[i];
            d[i] = (fabsf(dlf[i]) < 1e-8f) ? 0.0 : (double)dlf[i];
        }
        for (int i = 0; i < 3; ++i) {
            const double p = o[i] + d[i] * ts;
            int c = (int)floor(p);
            cur[i] = c < 0 ? 0 : (c > N - 1 ? N - 1 : c);
            if (d[i] > 0.0) {
                step[i] = 1;
                tDelta[i] = 1.0 / d[i];
                tNext[i] = ((double)(cur[i] + 1) - o[i]) / d[i];
            } else if (d[i] < 0.0) {
                step[i] = -1;
                tDelta[i] = -1.0 / d[i];
                tNext[i] = ((double)cur[i] - o[i]) / d[i];
            } else {
                step[i] = 0;
                tDelta[i] = INFINITY;
                tNext[i] = INFINITY;
            }
        }
        for (int i = 0; i < 3; ++i) {
            /* start within eps of a boundary: the previous voxel on that axis is a candidate */
            if (step[i] != 0) {
                const double tprev = tNext[i] - tDelta[i];
                if (fabs(tprev - ts) <= eps * fmax(1.0, fabs(ts))) {
                    int pc[3] = {cur[0], cur[1], cur[2]};
                    pc[i] -= step[i];
                    if (solid_at(grid, N, pc[0], pc[1], pc[2])) amb[r] = 1;
                }
            }
        }
        for (int guard = 0; guard < 4 * N + 8; ++guard) {
            if (cur[0] < 0 || cur[1] < 0 || cur[2] < 0 || cur[0] >= N || cur[1] >= N || cur[2] >= N) break;
            if (solid_at(grid, N, cur[0], cur[1], cur[2])) {
                double tn = -INFINITY;
                for (int i = 0; i < 3; ++i) {
                    if (d[i] == 0.0) continue;
                    const double t0 = ((double)cur[i] - o[i]) / d[i];
                    const double t1 = ((double)cur[i] + 1.0 - o[i]) / d[i];
                    const double a = t0 < t1 ? t0 : t1;
                    if (a > tn) tn = a;
                }
                if (tn < ts) tn = ts;
                vox[3 * r] = cur[0];
                vox[3 * r + 1] = cur[1];
                vox[3 * r + 2] = cur[2];
                tout[r] = tn;
                if (fabs(tn - te) <= eps * fmax(1.0, fabs(te))) amb[r] = 1;
                break;
            }
            int a = 0;
            if (tNext[1] < tNext[a]) a = 1;
            if (tNext[2] < tNext[a]) a = 2;
            const double tcross = tNext[a];
            if (tcross > te) {
                if (tcross - te <= eps * fmax(1.0, fabs(te))) {
                    int nc[3] = {cur[0], cur[1], cur[2]};
                    nc[a] += step[a];
                    if (solid_at(grid, N, nc[0], nc[1], nc[2])) amb[r] = 1;
                }
                break;
            }
            /* axes whose crossing is within eps of this one: an edge/corner tie */
            const double tol = eps * fmax(1.0, fabs(tcross));
            int tied = 0;
            for (int i = 0; i < 3; ++i)
                if (step[i] != 0 && tNext[i] - tcross <= tol) tied |= 1 << i;
            /* axes too close to parallel to step in FP32 still step here */
            for (int i = 0; i < 3; ++i)
                if ((tied >> i & 1) && fabsf(du[i]) < DIR_EPS) amb[r] = 1;
            if (tied & (tied - 1)) {
                for (int sub = 1; sub < 7; ++sub) {
                    if ((sub & tied) != sub || sub == tied) continue;
                    int nc[3] = {cur[0], cur[1], cur[2]};
                    for (int i = 0; i < 3; ++i)
                        if (sub >> i & 1) nc[i] += step[i];
                    if (solid_at(grid, N, nc[0], nc[1], nc[2])) amb[r] = 1;
                }
            }
            cur[a] += step[a];
            tNext[a] += tDelta[a];
        }
    }
}
