/*
 * vx_cpu_reference.c -- TEST INFRASTRUCTURE: a verbatim restatement of the
 * reference's CPU castRay for single-level worlds, to tie the oracle's
 * traversal (D1, vx_oracle.c) to the reference's own CPU semantics where those
 * are valid. Never linked into the product.
 *
 * Restated, in strict FP32 and the reference's evaluation order:
 *   castRayImpl                 libraries/SVO/src/SVOTraversal.cpp:516-691
 *   validateRayInput            :238-259 (glm::normalize as v * (1/sqrt(dot(v,v))))
 *   intersectAABB               :117-149
 *   initializeTraversalState    :261-280, selectInitialOctant :730-766
 *   fetchChildDescriptor        :282-293 (mirrorMask, SVOTypes.h:343-364)
 *   checkChildValidity          :295-326 (forced leaf at brickUserScale :304-308)
 *   executeAdvancePhase         :380-416
 *   executePopPhase at the root :418-431 (the exit test)
 *   computeRayCoefficients      :699-728, computeCorrectedTcMax :768-784,
 *   computeVoxelCorners         :786-794
 *   handleLeafHit               libraries/SVO/src/SVOBrickDDA.cpp:65-181
 *   traverseBrickAndReturnHit   :187-225
 *   traverseBrickView           :390-533 (occupancy: brickData word != 0, the
 *                               stand-in for Density > 0 read through the ECS)
 * Scope: worlds of at most 2 bricks per axis in voxel units (grid_max -
 * grid_min = N), the only worlds the reference's own tests build
 * (test_ray_casting_comprehensive.cpp:35-80, test_brick_traversal.cpp:28-30,
 * test_cornell_box.cpp:58). There every valid child of the root is a brick, so
 * the traversal stays at the root: PUSH and the integer POP below the root,
 * where the reference is defective, are never reached (a ray that would reach
 * them is reported as unsupported). getBrickViewByGrid and the leaf-to-brick
 * map are the root's leaf descriptors (one brick per grid cell).
 *
 * adjust: bits that replace one reference step each by the oracle's (D1)
 * choice, so the tests can attribute every disagreement with the oracle to a
 * named step (each bit's ray-level effect is checked against the FP64 brute
 * force); adjust = 0 is the reference as written:
 *   ORC_CPU_NO_UNIT_TMAX   t_max is not clamped to one unit cube
 *                          (castRayImpl :560-574 clamps it to 1.0, cutting rays
 *                          that cross more than one world edge of the cube)
 *   ORC_CPU_OCTANT_EPS     position-based first octant below t_min 1e-4
 *                          (ESVOTraversal.glsl:121) instead of 0.01 (:735)
 *   ORC_CPU_CLIP_BRICK     the brick span is clipped to the traced segment's
 *                          end and an empty span is no hit (traverseBrickView
 *                          tests the clamped start voxel even when the ray
 *                          misses the brick, SVOBrickDDA.cpp:404-414,455-468)
 *   ORC_CPU_LOOKUP_INSIDE  the brick under a leaf child is looked up 0.001
 *                          inside the child octant (handleLeafHit offsets its
 *                          lower corner along the ray, :94-101, which lands in
 *                          the neighbouring brick for rays going -axis)
 *
 * Records use the oracle's layout: flags = hit | entry axis << 1 | positive
 * normal << 3 | iterations << 16, t = the hit's tMin (world units), brick and
 * voxel of the hit, rgb decoded as the oracle decodes it; leaf_node = ~0u (the
 * reference's RayHit carries an entity, not a descriptor).
 */
#include "vx_oracle.h"

#include <float.h>
#include <math.h>
#include <string.h>

static inline float mn(float a, float b) { return (b < a) ? b : a; }  /* std::min, glm::min */
static inline float mx(float a, float b) { return (a < b) ? b : a; }  /* std::max, glm::max */
static inline float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); } /* glm::clamp */

typedef struct {
    float tx_coef, ty_coef, tz_coef, tx_bias, ty_bias, tz_bias;
    int octant_mask;
    float dir[3], norm_origin[3];
} coefs;

/* computeRayCoefficients, SVOTraversal.cpp:699-728 */
static coefs ray_coefficients(const float d[3], const float no[3])
{
    coefs c;
    const float eps = 1e-5f;
    float safe[3];
    for (int i = 0; i < 3; ++i) {
        c.dir[i] = d[i];
        c.norm_origin[i] = no[i];
        safe[i] = fabsf(d[i]) < eps ? copysignf(eps, d[i]) : d[i];
    }
    c.tx_coef = 1.0f / -fabsf(safe[0]);
    c.ty_coef = 1.0f / -fabsf(safe[1]);
    c.tz_coef = 1.0f / -fabsf(safe[2]);
    c.tx_bias = c.tx_coef * no[0];
    c.ty_bias = c.ty_coef * no[1];
    c.tz_bias = c.tz_coef * no[2];
    c.octant_mask = 7;
    if (d[0] > 0.0f) { c.octant_mask ^= 1; c.tx_bias = 3.0f * c.tx_coef - c.tx_bias; }
    if (d[1] > 0.0f) { c.octant_mask ^= 2; c.ty_bias = 3.0f * c.ty_coef - c.ty_bias; }
    if (d[2] > 0.0f) { c.octant_mask ^= 4; c.tz_bias = 3.0f * c.tz_coef - c.tz_bias; }
    return c;
}

/* computeVoxelCorners :786-794 */
static void corners(const float pos[3], const coefs* c, float* tx, float* ty, float* tz)
{
    *tx = pos[0] * c->tx_coef - c->tx_bias;
    *ty = pos[1] * c->ty_coef - c->ty_bias;
    *tz = pos[2] * c->tz_coef - c->tz_bias;
}

/* computeCorrectedTcMax :768-784 (std::min over an initializer list reduces left to right) */
static float corrected_tc_max(float tx, float ty, float tz, const float d[3], float t_max)
{
    const float thr = 1000.0f, eps = 1e-5f;
    const float vx = (fabsf(d[0]) >= eps && fabsf(tx) < thr) ? tx : t_max;
    const float vy = (fabsf(d[1]) >= eps && fabsf(ty) < thr) ? ty : t_max;
    const float vz = (fabsf(d[2]) >= eps && fabsf(tz) < thr) ? tz : t_max;
    return mn(mn(vx, vy), vz);
}

/* mirrorMask, SVOTypes.h:343-364 */
static uint32_t mirror_mask(uint32_t mask, int octant_mask)
{
    if (octant_mask == 7) return mask & 0xFFu;
    const int flip = (~octant_mask) & 7;
    uint32_t r = 0;
    for (int i = 0; i < 8; ++i)
        if (mask & (1u << i)) r |= 1u << (i ^ flip);
    return r;
}

/* The brick of root child octant `oct` (the grid cell bx | by << 1 | bz << 2 of a
 * single-level world), or -1: getBrickViewByGrid / getBrickView. */
static int64_t root_brick(const orc_scene* s, int oct)
{
    const uint32_t w0 = s->nodes[0], w1 = s->nodes[1];
    const uint32_t valid = (w0 >> 16) & 0xFFu, leaf = (w0 >> 24) & 0xFFu;
    if (!((valid >> oct) & 1u) || !((leaf >> oct) & 1u)) return -1;
    const uint32_t cp = (w0 & 0x8000u) ? (w1 & 0xFFFFFFu) : (w0 & 0x7FFFu);
    const uint32_t below = (1u << oct) - 1u;
    const uint64_t li = (uint64_t)cp + (uint64_t)__builtin_popcount(valid & ~leaf & 0xFFu) +
                        (uint64_t)__builtin_popcount(valid & leaf & below);
    if (li >= s->node_count) return -1;
    const uint32_t b = s->nodes[2 * li + 1] & 0xFFFFFFu;
    return (b == 0xFFFFFFu || b >= s->brick_count) ? -1 : (int64_t)b;
}

static int64_t brick_by_grid(const orc_scene* s, const int bi[3])
{
    return root_brick(s, bi[0] | (bi[1] << 1) | (bi[2] << 2));
}

static void set_miss(orc_hit* h, uint32_t iters)
{
    memset(h, 0, sizeof *h);
    h->flags = iters << 16;
    h->t = INFINITY;
    h->leaf_node = h->brick = h->voxel = 0xFFFFFFFFu;
}

/* traverseBrickView, SVOBrickDDA.cpp:390-533, with brickVoxelSize 1 and the
 * brick's grid origin bmin. Returns 1 on a hit (voxel, t, entry axis and sign). */
static int brick_view_dda(const orc_scene* s, uint32_t brick, const float bmin[3], const float o[3],
                          const float d[3], float tMin, float tMax, uint32_t* voxel, float* t_hit, int* axis,
                          int* sign_pos)
{
    const int brickN = 8;
    float entry[3], tDelta[3], tNext[3];
    int cur[3], step[3];
    for (int i = 0; i < 3; ++i) {
        entry[i] = o[i] + d[i] * tMin;
        const float local = (entry[i] - bmin[i]) / 1.0f;
        int c = (int)floorf(local);
        cur[i] = c < 0 ? 0 : (c > brickN - 1 ? brickN - 1 : c);
    }
    for (int i = 0; i < 3; ++i) {
        if (fabsf(d[i]) < 1e-8f) {
            step[i] = 0;
            tDelta[i] = FLT_MAX;
            tNext[i] = FLT_MAX;
        } else {
            step[i] = d[i] > 0.0f ? 1 : -1;
            tDelta[i] = 1.0f / fabsf(d[i]);
            if (d[i] > 0.0f) {
                const float nb = bmin[i] + (float)(cur[i] + 1) * 1.0f;
                const float dist = nb - entry[i];
                tNext[i] = tMin + dist / d[i];
            } else {
                const float nb = bmin[i] + (float)cur[i] * 1.0f;
                const float dist = entry[i] - nb;
                tNext[i] = tMin + dist / fabsf(d[i]);
            }
        }
    }
    for (int steps = 0; steps < brickN * 3;) {
        ++steps;
        if (cur[0] < 0 || cur[0] >= brickN || cur[1] < 0 || cur[1] >= brickN || cur[2] < 0 || cur[2] >= brickN)
            return 0;
        const uint32_t lin = (uint32_t)(cur[0] + 8 * cur[1] + 64 * cur[2]);
        if (s->brick_data[(uint64_t)brick * 512u + lin] != 0u) {
            float tn[3];
            for (int i = 0; i < 3; ++i) {
                const float vmin = bmin[i] + (float)cur[i] * 1.0f;
                const float vmax = vmin + 1.0f;
                float t0, t1;
                if (fabsf(d[i]) < 1e-8f) {
                    t0 = -INFINITY;
                    t1 = INFINITY;
                } else {
                    t0 = (vmin - o[i]) / d[i];
                    t1 = (vmax - o[i]) / d[i];
                }
                tn[i] = mn(t0, t1);
            }
            float ht = mx(mx(tn[0], tn[1]), tn[2]);
            ht = mx(ht, 0.0f);
            int a;
            if (tn[0] >= tn[1] && tn[0] >= tn[2]) a = 0;
            else if (tn[1] >= tn[2]) a = 1;
            else a = 2;
            *voxel = lin;
            *t_hit = ht;
            *axis = a;
            *sign_pos = d[a] > 0.0f ? 0 : 1; /* entryNormal = rayDir > 0 ? -1 : +1 */
            return 1;
        }
        if (tNext[0] < tNext[1] && tNext[0] < tNext[2]) {
            if (tNext[0] > tMax) return 0;
            cur[0] += step[0];
            tNext[0] += tDelta[0];
        } else if (tNext[1] < tNext[2]) {
            if (tNext[1] > tMax) return 0;
            cur[1] += step[1];
            tNext[1] += tDelta[1];
        } else {
            if (tNext[2] > tMax) return 0;
            cur[2] += step[2];
            tNext[2] += tDelta[2];
        }
    }
    return 0;
}

/* 1 hit, 0 miss, -1 the ray reaches a part the restatement does not cover. */
static int cast_one(const orc_scene* s, const orc_ray* r, orc_hit* h, uint32_t adjust)
{
    const float* wmin = s->grid_min;
    const float* wmax = s->grid_max;
    const int B = s->bricks_per_axis, maxLevels = s->max_levels;
    set_miss(h, 0);
    /* validateRayInput */
    float d[3];
    const float dot = r->d[0] * r->d[0] + r->d[1] * r->d[1] + r->d[2] * r->d[2];
    const float inv = 1.0f / sqrtf(dot);
    for (int i = 0; i < 3; ++i) d[i] = r->d[i] * inv;
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(len >= 1e-6f)) return 0; /* rayLength < 1e-6 (and NaN) */
    for (int i = 0; i < 3; ++i)
        if (isnan(r->o[i]) || isnan(d[i]) || isinf(r->o[i]) || isinf(d[i])) return 0;
    /* isPointInsideAABB, intersectAABB */
    const int inside = r->o[0] >= wmin[0] && r->o[0] <= wmax[0] && r->o[1] >= wmin[1] && r->o[1] <= wmax[1] &&
                       r->o[2] >= wmin[2] && r->o[2] <= wmax[2];
    float invd[3];
    for (int i = 0; i < 3; ++i) {
        if (fabsf(d[i]) < 1e-8f) {
            if (r->o[i] < wmin[i] || r->o[i] > wmax[i]) return 0;
            invd[i] = d[i] >= 0.0f ? 1e20f : -1e20f;
        } else {
            invd[i] = 1.0f / d[i];
        }
    }
    float tn[3], tf[3];
    for (int i = 0; i < 3; ++i) {
        const float t0 = (wmin[i] - r->o[i]) * invd[i], t1 = (wmax[i] - r->o[i]) * invd[i];
        tn[i] = mn(t0, t1);
        tf[i] = mx(t0, t1);
    }
    float tEntry = mx(mx(tn[0], tn[1]), tn[2]);
    float tExit = mn(mn(tf[0], tf[1]), tf[2]);
    if (!(tEntry <= tExit && tExit >= 0.0f)) return 0;
    tEntry = mx(tEntry, r->tmin);
    tExit = mn(tExit, r->tmax);
    if (tEntry >= tExit || tExit < 0.0f) return 0;
    const float tRayStart = inside ? 0.0f : mx(0.0f, tEntry);
    float ws[3], no[3];
    for (int i = 0; i < 3; ++i) {
        ws[i] = wmax[i] - wmin[i];
        const float entry = r->o[i] + d[i] * tRayStart;
        no[i] = (entry - wmin[i]) / ws[i] + 1.0f;
    }
    const coefs c = ray_coefficients(d, no);
    float t_min, t_max;
    if (inside) {
        t_min = 0.0f;
        t_max = mn(mn(c.tx_coef - c.tx_bias, c.ty_coef - c.ty_bias), c.tz_coef - c.tz_bias);
        if (!(adjust & ORC_CPU_NO_UNIT_TMAX)) t_max = mn(t_max, 1.0f);
    } else {
        t_min = mx(mx(2.0f * c.tx_coef - c.tx_bias, 2.0f * c.ty_coef - c.ty_bias), 2.0f * c.tz_coef - c.tz_bias);
        t_max = mn(mn(c.tx_coef - c.tx_bias, c.ty_coef - c.ty_bias), c.tz_coef - c.tz_bias);
        t_min = mx(t_min, 0.0f);
        if (!(adjust & ORC_CPU_NO_UNIT_TMAX)) t_max = mn(t_max, 1.0f);
    }
    /* initializeTraversalState + selectInitialOctant */
    int scale = 22, idx = 0, iter = 0;
    float pos[3] = {1.0f, 1.0f, 1.0f}, scale_exp2 = 0.5f;
    {
        const int use_pos = t_min < ((adjust & ORC_CPU_OCTANT_EPS) ? 1e-4f : 0.01f);
        const float m[3] = {(c.octant_mask & 1) ? no[0] : (3.0f - no[0]), (c.octant_mask & 2) ? no[1] : (3.0f - no[1]),
                            (c.octant_mask & 4) ? no[2] : (3.0f - no[2])};
        const float coef[3] = {c.tx_coef, c.ty_coef, c.tz_coef}, bias[3] = {c.tx_bias, c.ty_bias, c.tz_bias};
        for (int i = 0; i < 3; ++i) {
            if (fabsf(d[i]) < 1e-5f || use_pos) {
                if (m[i] >= 1.5f) { idx |= 1 << i; pos[i] = 1.5f; }
            } else if (1.5f * coef[i] - bias[i] > t_min) {
                idx ^= 1 << i;
                pos[i] = 1.5f;
            }
        }
    }
    const int minScale = 22 - maxLevels + 1;
    const uint32_t mvalid = mirror_mask((s->nodes[0] >> 16) & 0xFFu, c.octant_mask);
    const uint32_t mleaf = mirror_mask((s->nodes[0] >> 24) & 0xFFu, c.octant_mask);
    while (scale >= minScale && scale <= 22 && iter < 500) {
        ++iter;
        /* checkChildValidity */
        const int valid = (mvalid >> idx) & 1u;
        int is_leaf = (mleaf >> idx) & 1u;
        if ((scale - minScale) == maxLevels - 3 && valid) is_leaf = 1;
        float tv_max = 0.0f;
        int process = 0;
        if (valid && !(t_min > t_max)) {
            float tx, ty, tz;
            corners(pos, &c, &tx, &ty, &tz);
            tv_max = mn(t_max, corrected_tc_max(tx, ty, tz, d, t_max));
            process = t_min <= tv_max;
        }
        if (process) {
            if (!is_leaf) return -1; /* PUSH: not a single-level world */
            /* handleLeafHit */
            float lp[3], inside_oct[3], rpl[3];
            const float off = 0.001f;
            for (int i = 0; i < 3; ++i) {
                lp[i] = (c.octant_mask & (1 << i)) ? pos[i] : (3.0f - scale_exp2 - pos[i]);
                const float ln = lp[i] - 1.0f;
                inside_oct[i] = ln + ((adjust & ORC_CPU_LOOKUP_INSIDE) || d[i] > 0.0f ? off : -off);
                const float rp = r->o[i] + d[i] * mx(tEntry, 0.0f);
                rpl[i] = clampf((rp - wmin[i]) / ws[i], 0.001f, 0.999f);
            }
            for (int i = 0; i < 3; ++i) {
                if (fabsf(d[i]) < 1e-5f) inside_oct[i] = rpl[i];
                inside_oct[i] = clampf(inside_oct[i], 0.001f, 0.999f);
            }
            int bi[3];
            for (int i = 0; i < 3; ++i) {
                const float hp = clampf(inside_oct[i] * ws[i], 0.0f, ws[i] - 0.001f);
                int b = (int)(hp / 8.0f);
                bi[i] = b < 0 ? 0 : (b > B - 1 ? B - 1 : b);
            }
            int64_t brick = brick_by_grid(s, bi);
            if (brick < 0) {
                for (int i = 0; i < 3; ++i) {
                    float e = (r->o[i] + d[i] * mx(tEntry, 0.0f)) - wmin[i];
                    e = e + d[i] * 0.01f;
                    e = clampf(e, 0.0f, ws[i] - 0.001f);
                    int b = (int)(e / 8.0f);
                    bi[i] = b < 0 ? 0 : (b > B - 1 ? B - 1 : b);
                }
                brick = brick_by_grid(s, bi);
            }
            if (brick < 0) {
                const int local = idx ^ ((~c.octant_mask) & 7);
                brick = root_brick(s, local);
                for (int i = 0; i < 3; ++i) bi[i] = (local >> i) & 1;
            }
            if (brick >= 0) {
                /* traverseBrickAndReturnHit: the ray in local grid space (worldToLocal = translate(-worldMin)) */
                float lo[3], bmin[3], bmax[3], id[3];
                for (int i = 0; i < 3; ++i) {
                    lo[i] = r->o[i] - wmin[i];
                    bmin[i] = (float)(bi[i] * 8);
                    bmax[i] = bmin[i] + (float)8 * 1.0f;
                    id[i] = fabsf(d[i]) < 1e-8f ? (d[i] >= 0.0f ? 1e8f : -1e8f) : 1.0f / d[i];
                }
                float bn[3], bf[3];
                for (int i = 0; i < 3; ++i) {
                    const float t0 = (bmin[i] - lo[i]) * id[i], t1 = (bmax[i] - lo[i]) * id[i];
                    bn[i] = mn(t0, t1);
                    bf[i] = mx(t0, t1);
                }
                float btmin = mx(mx(bn[0], bn[1]), bn[2]);
                float btmax = mn(mn(bf[0], bf[1]), bf[2]);
                btmin = mx(btmin, tEntry);
                if (adjust & ORC_CPU_CLIP_BRICK) btmax = mn(btmax, tExit);
                uint32_t voxel;
                float th;
                int axis, sp;
                if (!((adjust & ORC_CPU_CLIP_BRICK) && btmin > btmax) &&
                    brick_view_dda(s, (uint32_t)brick, bmin, lo, d, btmin, btmax, &voxel, &th, &axis, &sp)) {
                    h->flags = 1u | ((uint32_t)axis << 1) | ((uint32_t)sp << 3) | ((uint32_t)iter << 16);
                    h->t = th;
                    h->leaf_node = 0xFFFFFFFFu;
                    h->brick = (uint32_t)brick;
                    h->voxel = voxel;
                    orc_voxel_color(s, brick, voxel, s->brick_data[(uint64_t)brick * 512u + voxel], h->rgb);
                    return 1;
                }
            }
            t_min = tv_max; /* then ADVANCE */
        }
        /* executeAdvancePhase */
        float tx, ty, tz;
        corners(pos, &c, &tx, &ty, &tz);
        const int cx = fabsf(d[0]) >= 1e-5f, cy = fabsf(d[1]) >= 1e-5f, cz = fabsf(d[2]) >= 1e-5f;
        float tcm = corrected_tc_max(tx, ty, tz, d, t_max);
        if (tcm == FLT_MAX)
            tcm = mx(mx(cx ? tx : -FLT_MAX, cy ? ty : -FLT_MAX), cz ? tz : -FLT_MAX);
        int step_mask = 0;
        if (cx && tx <= tcm) { step_mask ^= 1; pos[0] -= scale_exp2; }
        if (cy && ty <= tcm) { step_mask ^= 2; pos[1] -= scale_exp2; }
        if (cz && tz <= tcm) { step_mask ^= 4; pos[2] -= scale_exp2; }
        t_min = mx(tcm, 0.0f);
        idx ^= step_mask;
        if ((idx & step_mask) != 0) {
            /* executePopPhase at the root: leave the octree or go on */
            if (scale != 22) return -1;
            if (t_min > t_max || pos[0] < 1.0f || pos[0] >= 2.0f || pos[1] < 1.0f || pos[1] >= 2.0f ||
                pos[2] < 1.0f || pos[2] >= 2.0f)
                break;
        }
    }
    h->flags = (uint32_t)iter << 16;
    return 0;
}

int orc_trace_cpu_reference(const orc_scene* s, const orc_ray* rays, uint64_t n, uint32_t adjust, orc_hit* out,
                            int32_t* status)
{
    const int N = s->bricks_per_axis * 8;
    if (s->bricks_per_axis > 2 || s->node_count == 0) return -1;
    for (int i = 0; i < 3; ++i)
        if (s->grid_max[i] - s->grid_min[i] != (float)N) return -1;
    for (uint64_t k = 0; k < n; ++k) status[k] = cast_one(s, &rays[k], &out[k], adjust);
    return 0;
}
