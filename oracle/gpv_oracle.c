/*
 * gpv_oracle.c -- CPU restatement of GPView's two-level hybrid voxelizer.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, and the CPU
 * baseline that is timed beside it.  Nothing in the product links it.
 *
 * Float semantics: every expression keeps the reference's operand order and
 * its float/double promotions (the reference mixes double literals such as
 * 0.5, 1.0 and 1e-6 into float expressions).  Compile with
 * -ffp-contract=off and without -ffast-math.
 *
 * Two deliberate departures, both where the reference's behaviour is
 * undefined rather than a result:
 *   - index-AABB lower bounds are clamped at 0 (the reference indexes out of
 *     bounds for a vertex below the bbox; our bbox always contains the mesh);
 *   - Level2 storage starts zeroed for every boundary voxel (the reference
 *     memsets only the first Level2 block, src/Object.cpp:3284).
 */
#include "gpv_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small growable arrays                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int64_t *a; int64_t n, cap; } vec64;

static void v64_push(vec64 *v, int64_t x)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 1024;
        v->a = (int64_t *)realloc(v->a, (size_t)v->cap * sizeof(int64_t));
        if (!v->a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    }
    v->a[v->n++] = x;
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz); abort(); }
    return p;
}

/* ------------------------------------------------------------------------ */
/* SAT triangle/box overlap: src/TriBoxIntersection.cpp:64-210               */
/* ------------------------------------------------------------------------ */

/* planeBoxOverlap, src/TriBoxIntersection.cpp:64-86 */
static int plane_box_overlap(const float nrm[3], const float vert[3], const float box[3])
{
    float vmin[3], vmax[3];
    for (int q = 0; q < 3; q++) {
        float v = vert[q];
        if (nrm[q] > 0.0f) {
            vmin[q] = -box[q] - v;
            vmax[q] = box[q] - v;
        } else {
            vmin[q] = box[q] - v;
            vmax[q] = -box[q] - v;
        }
    }
    if (nrm[0] * vmin[0] + nrm[1] * vmin[1] + nrm[2] * vmin[2] > 0.0f) return 0;
    if (nrm[0] * vmax[0] + nrm[1] * vmax[1] + nrm[2] * vmax[2] >= 0.0f) return 1;
    return 0;
}

/* One edge-axis test (the AXISTEST_* macros, src/TriBoxIntersection.cpp:90-133):
 * separated when the projected interval [min(pa,pb), max(pa,pb)] misses
 * [-rad, rad] strictly.  (pa, pb) is the macro's compare order: (p0, p2) for
 * X01/Y02, (p0, p1) for X2/Y1/Z0, (p2, p1) for Z12 (:124); it decides which
 * of a NaN pair becomes min. */
static int separated(float pa, float pb, float rad)
{
    float mn, mx;
    if (pa < pb) { mn = pa; mx = pb; } else { mn = pb; mx = pa; }
    return (mn > rad || mx < -rad);
}

/* TriBoxOverlap, src/TriBoxIntersection.cpp:135-210 */
int orc_tri_box_overlap(const float c[3], const float h[3], const float tv[9])
{
    float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3], n[3];
    float fx, fy, fz;
    for (int q = 0; q < 3; q++) {
        v0[q] = tv[q] - c[q];
        v1[q] = tv[3 + q] - c[q];
        v2[q] = tv[6 + q] - c[q];
    }
    for (int q = 0; q < 3; q++) {
        e0[q] = v1[q] - v0[q];
        e1[q] = v2[q] - v1[q];
        e2[q] = v0[q] - v2[q];
    }
    /* edge e0: X01, Y02, Z12 (:163-168) */
    fx = fabsf(e0[0]); fy = fabsf(e0[1]); fz = fabsf(e0[2]);
    if (separated(e0[2] * v0[1] - e0[1] * v0[2], e0[2] * v2[1] - e0[1] * v2[2],
                  fz * h[1] + fy * h[2])) return 0;
    if (separated(-e0[2] * v0[0] + e0[0] * v0[2], -e0[2] * v2[0] + e0[0] * v2[2],
                  fz * h[0] + fx * h[2])) return 0;
    if (separated(e0[1] * v2[0] - e0[0] * v2[1], e0[1] * v1[0] - e0[0] * v1[1],
                  fy * h[0] + fx * h[1])) return 0;
    /* edge e1: X01, Y02, Z0 (:170-175) */
    fx = fabsf(e1[0]); fy = fabsf(e1[1]); fz = fabsf(e1[2]);
    if (separated(e1[2] * v0[1] - e1[1] * v0[2], e1[2] * v2[1] - e1[1] * v2[2],
                  fz * h[1] + fy * h[2])) return 0;
    if (separated(-e1[2] * v0[0] + e1[0] * v0[2], -e1[2] * v2[0] + e1[0] * v2[2],
                  fz * h[0] + fx * h[2])) return 0;
    if (separated(e1[1] * v0[0] - e1[0] * v0[1], e1[1] * v1[0] - e1[0] * v1[1],
                  fy * h[0] + fx * h[1])) return 0;
    /* edge e2: X2, Y1, Z12 (:177-182) */
    fx = fabsf(e2[0]); fy = fabsf(e2[1]); fz = fabsf(e2[2]);
    if (separated(e2[2] * v0[1] - e2[1] * v0[2], e2[2] * v1[1] - e2[1] * v1[2],
                  fz * h[1] + fy * h[2])) return 0;
    if (separated(-e2[2] * v0[0] + e2[0] * v0[2], -e2[2] * v1[0] + e2[0] * v1[2],
                  fz * h[0] + fx * h[2])) return 0;
    if (separated(e2[1] * v2[0] - e2[0] * v2[1], e2[1] * v1[0] - e2[0] * v1[1],
                  fy * h[0] + fx * h[1])) return 0;
    /* AABB axes, FINDMINMAX (:191-200) */
    for (int q = 0; q < 3; q++) {
        float mn = v0[q], mx = v0[q];
        if (v1[q] < mn) mn = v1[q];
        if (v1[q] > mx) mx = v1[q];
        if (v2[q] < mn) mn = v2[q];
        if (v2[q] > mx) mx = v2[q];
        if (mn > h[q] || mx < -h[q]) return 0;
    }
    /* plane of the triangle, n = e0 x e1 (:205-207, CROSS :45-48) */
    n[0] = e0[1] * e1[2] - e0[2] * e1[1];
    n[1] = e0[2] * e1[0] - e0[0] * e1[2];
    n[2] = e0[0] * e1[1] - e0[1] * e1[0];
    return plane_box_overlap(n, v0, h);
}

/* ------------------------------------------------------------------------ */
/* The SAT factored by the center components it reads (fast mode only)      */
/* ------------------------------------------------------------------------ */
/* orc_tri_box_overlap is the conjunction of 13 side-effect-free predicates.
 * With v_k = tri_k - c taken component by component and e_k the differences
 * of the v_k, component q of every v_k and e_k depends on c[q] alone, so:
 *   the AABB test of axis q reads c[q] only;
 *   the three edge tests along e x X read components 1 and 2 (c.y, c.z),
 *   along e x Y components 0 and 2, along e x Z components 0 and 1;
 *   only the plane test reads all three.
 * The helpers below evaluate exactly those float expressions (same operands,
 * same order) for one or two center components, so a grid of centers can be
 * screened by 1-D and 2-D tables; a center that passes every screen has
 * passed 12 of the 13 predicates and is decided by the plane test alone.
 * No tolerance or error bound is involved. */
static int aabb_axis_ok(const float *tv, int q, float c, float h)
{
    float a = tv[q] - c, b = tv[3 + q] - c, d = tv[6 + q] - c;
    float mn = a, mx = a;
    if (b < mn) mn = b;
    if (b > mx) mx = b;
    if (d < mn) mn = d;
    if (d > mx) mx = d;
    return !(mn > h || mx < -h);
}

/* the three tests along e_k x X (AXISTEST_X01 e0, X01 e1, X2 e2) */
static int edges_x_ok(const float *tv, float cy, float cz, const float h[3])
{
    float v0y = tv[1] - cy, v1y = tv[4] - cy, v2y = tv[7] - cy;
    float v0z = tv[2] - cz, v1z = tv[5] - cz, v2z = tv[8] - cz;
    float e0y = v1y - v0y, e1y = v2y - v1y, e2y = v0y - v2y;
    float e0z = v1z - v0z, e1z = v2z - v1z, e2z = v0z - v2z;
    if (separated(e0z * v0y - e0y * v0z, e0z * v2y - e0y * v2z, fabsf(e0z) * h[1] + fabsf(e0y) * h[2])) return 0;
    if (separated(e1z * v0y - e1y * v0z, e1z * v2y - e1y * v2z, fabsf(e1z) * h[1] + fabsf(e1y) * h[2])) return 0;
    if (separated(e2z * v0y - e2y * v0z, e2z * v1y - e2y * v1z, fabsf(e2z) * h[1] + fabsf(e2y) * h[2])) return 0;
    return 1;
}

/* the three tests along e_k x Y (AXISTEST_Y02 e0, Y02 e1, Y1 e2) */
static int edges_y_ok(const float *tv, float cx, float cz, const float h[3])
{
    float v0x = tv[0] - cx, v1x = tv[3] - cx, v2x = tv[6] - cx;
    float v0z = tv[2] - cz, v1z = tv[5] - cz, v2z = tv[8] - cz;
    float e0x = v1x - v0x, e1x = v2x - v1x, e2x = v0x - v2x;
    float e0z = v1z - v0z, e1z = v2z - v1z, e2z = v0z - v2z;
    if (separated(-e0z * v0x + e0x * v0z, -e0z * v2x + e0x * v2z, fabsf(e0z) * h[0] + fabsf(e0x) * h[2])) return 0;
    if (separated(-e1z * v0x + e1x * v0z, -e1z * v2x + e1x * v2z, fabsf(e1z) * h[0] + fabsf(e1x) * h[2])) return 0;
    if (separated(-e2z * v0x + e2x * v0z, -e2z * v1x + e2x * v1z, fabsf(e2z) * h[0] + fabsf(e2x) * h[2])) return 0;
    return 1;
}

/* the three tests along e_k x Z (AXISTEST_Z12 e0, Z0 e1, Z12 e2) */
static int edges_z_ok(const float *tv, float cx, float cy, const float h[3])
{
    float v0x = tv[0] - cx, v1x = tv[3] - cx, v2x = tv[6] - cx;
    float v0y = tv[1] - cy, v1y = tv[4] - cy, v2y = tv[7] - cy;
    float e0x = v1x - v0x, e1x = v2x - v1x, e2x = v0x - v2x;
    float e0y = v1y - v0y, e1y = v2y - v1y, e2y = v0y - v2y;
    if (separated(e0y * v2x - e0x * v2y, e0y * v1x - e0x * v1y, fabsf(e0y) * h[0] + fabsf(e0x) * h[1])) return 0;
    if (separated(e1y * v0x - e1x * v0y, e1y * v1x - e1x * v1y, fabsf(e1y) * h[0] + fabsf(e1x) * h[1])) return 0;
    if (separated(e2y * v2x - e2x * v2y, e2y * v1x - e2x * v1y, fabsf(e2y) * h[0] + fabsf(e2x) * h[1])) return 0;
    return 1;
}

/* the plane predicate alone: n = e0 x e1, planeBoxOverlap(n, v0, h) */
static int plane_ok(const float *tv, const float c[3], const float h[3])
{
    float v0[3], v1[3], v2[3], e0[3], e1[3], n[3];
    for (int q = 0; q < 3; q++) {
        v0[q] = tv[q] - c[q];
        v1[q] = tv[3 + q] - c[q];
        v2[q] = tv[6 + q] - c[q];
    }
    for (int q = 0; q < 3; q++) {
        e0[q] = v1[q] - v0[q];
        e1[q] = v2[q] - v1[q];
    }
    n[0] = e0[1] * e1[2] - e0[2] * e1[1];
    n[1] = e0[2] * e1[0] - e0[0] * e1[2];
    n[2] = e0[0] * e1[1] - e0[1] * e1[0];
    return plane_box_overlap(n, v0, h);
}

/* The factored form on one case (for the tests: it must equal
 * orc_tri_box_overlap everywhere). */
int orc_tri_box_factored(const float c[3], const float h[3], const float tv[9])
{
    for (int q = 0; q < 3; q++)
        if (!aabb_axis_ok(tv, q, c[q], h[q])) return 0;
    if (!edges_z_ok(tv, c[0], c[1], h) || !edges_x_ok(tv, c[1], c[2], h) || !edges_y_ok(tv, c[0], c[2], h))
        return 0;
    return plane_ok(tv, c, h);
}

void orc_tri_box_factored_batch(int64_t n, const float *c, const float *h, const float *tv, uint8_t *res)
{
    for (int64_t i = 0; i < n; i++) res[i] = (uint8_t)orc_tri_box_factored(c + 3 * i, h + 3 * i, tv + 9 * i);
}

/* Per-thread tables of the sweep. */
typedef struct { uint8_t *a; size_t cap; } bytes_t;

static uint8_t *bytes_get(bytes_t *b, size_t n)
{
    if (n > b->cap) {
        b->a = (uint8_t *)realloc(b->a, n);
        if (!b->a) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
        b->cap = n;
    }
    return b->a;
}

/* Every (p, q, s) of the nx x ny x nz center grid cx[p], cy[q], cz[s] whose
 * box overlaps triangle tv: the literal triple loop over
 * orc_tri_box_overlap, screened by the factored predicates.  The plane
 * tables are built only over the indices whose AABB test passes.
 * emit(ud, p, q, s) is called for each overlap. */
static void sat_sweep(const float *tv, const float h[3], const float *cx, int nx, const float *cy, int ny,
                      const float *cz, int nz, bytes_t *tmp, void (*emit)(void *, int, int, int), void *ud)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return;
    int *ix = (int *)bytes_get(tmp, sizeof(int) * ((size_t)nx + ny + nz));
    int *iy = ix + nx, *iz = iy + ny;
    int mx = 0, my = 0, mz = 0;
    for (int p = 0; p < nx; p++)
        if (aabb_axis_ok(tv, 0, cx[p], h[0])) ix[mx++] = p;
    if (!mx) return;
    for (int q = 0; q < ny; q++)
        if (aabb_axis_ok(tv, 1, cy[q], h[1])) iy[my++] = q;
    if (!my) return;
    for (int k = 0; k < nz; k++)
        if (aabb_axis_ok(tv, 2, cz[k], h[2])) iz[mz++] = k;
    if (!mz) return;
    size_t need = sizeof(int) * ((size_t)nx + ny + nz) + (size_t)mx * my + (size_t)my * mz + (size_t)mx * mz;
    ix = (int *)bytes_get(tmp, need);   /* may move: re-derive the views */
    iy = ix + nx;
    iz = iy + ny;
    uint8_t *txy = (uint8_t *)(iz + nz), *tyz = txy + (size_t)mx * my, *txz = tyz + (size_t)my * mz;
    int any = 0;
    for (int a = 0; a < mx; a++)
        for (int b = 0; b < my; b++)
            any |= (txy[(size_t)a * my + b] = (uint8_t)edges_z_ok(tv, cx[ix[a]], cy[iy[b]], h));
    if (!any) return;
    any = 0;
    for (int b = 0; b < my; b++)
        for (int k = 0; k < mz; k++)
            any |= (tyz[(size_t)b * mz + k] = (uint8_t)edges_x_ok(tv, cy[iy[b]], cz[iz[k]], h));
    if (!any) return;
    for (int a = 0; a < mx; a++)
        for (int k = 0; k < mz; k++)
            txz[(size_t)a * mz + k] = (uint8_t)edges_y_ok(tv, cx[ix[a]], cz[iz[k]], h);
    for (int a = 0; a < mx; a++)
        for (int b = 0; b < my; b++) {
            if (!txy[(size_t)a * my + b]) continue;
            for (int k = 0; k < mz; k++) {
                if (!tyz[(size_t)b * mz + k] || !txz[(size_t)a * mz + k]) continue;
                float ctr[3] = {cx[ix[a]], cy[iy[b]], cz[iz[k]]};
                if (plane_ok(tv, ctr, h)) emit(ud, ix[a], iy[b], iz[k]);
            }
        }
}

/* ------------------------------------------------------------------------ */
/* Moller-Trumbore: src/TriRayIntersection.cpp:78-131 (EPSILON :38 is the   */
/* double 1e-6; det and t are promoted to double for the compares)          */
/* ------------------------------------------------------------------------ */
#define MT_EPS 0.000001

int orc_tri_ray(const float V1[3], const float V2[3], const float V3[3],
                const float O[3], const float D[3], float *t_out)
{
    float e1[3], e2[3], P[3], Q[3], T[3];
    float det, inv_det, u, v, t;
    for (int q = 0; q < 3; q++) { e1[q] = V2[q] - V1[q]; e2[q] = V3[q] - V1[q]; }
    P[0] = D[1] * e2[2] - D[2] * e2[1];
    P[1] = D[2] * e2[0] - D[0] * e2[2];
    P[2] = D[0] * e2[1] - D[1] * e2[0];
    det = e1[0] * P[0] + e1[1] * P[1] + e1[2] * P[2];
    if ((double)det > -MT_EPS && (double)det < MT_EPS) return 0;
    inv_det = 1.f / det;
    for (int q = 0; q < 3; q++) T[q] = O[q] - V1[q];
    u = (T[0] * P[0] + T[1] * P[1] + T[2] * P[2]) * inv_det;
    if (u < 0.f || u > 1.f) return 0;
    Q[0] = T[1] * e1[2] - T[2] * e1[1];
    Q[1] = T[2] * e1[0] - T[0] * e1[2];
    Q[2] = T[0] * e1[1] - T[1] * e1[0];
    v = (D[0] * Q[0] + D[1] * Q[1] + D[2] * Q[2]) * inv_det;
    if (v < 0.f || u + v > 1.f) return 0;
    t = (e2[0] * Q[0] + e2[1] * Q[1] + e2[2] * Q[2]) * inv_det;
    if ((double)t > MT_EPS) {
        if (t_out) *t_out = t;
        return 1;
    }
    return 0;
}

/* Column factorisation of the +z ray (D = (0,0,1)).  For that D every term
 * in P, det, u and v that involves O.z is multiplied by an exact zero, so the
 * det/u/v decision is a function of (O.x, O.y) alone and only t needs O.z.
 * mt_xy/mt_t evaluate the same float expressions as orc_tri_ray, split at
 * that seam.  Equality with orc_tri_ray is asserted by the tests. */
typedef struct {
    float v1[3], e1[3], e2[3];
    float inv_det;
    float tx, ty;           /* T.x, T.y for the column */
} mt_col;

static const float RAY_D[3] = {0.0f, 0.0f, 1.0f};

/* the column-independent half: edges, P = D x e2, det and its reciprocal;
 * returns 0 when the det test rejects every ray */
typedef struct { float v1[3], e1[3], e2[3], P[3], inv_det; } mt_tri;

static int mt_setup(const float *tv, mt_tri *m)
{
    float *e1 = m->e1, *e2 = m->e2, *P = m->P;
    for (int q = 0; q < 3; q++) { m->v1[q] = tv[q]; e1[q] = tv[3 + q] - tv[q]; e2[q] = tv[6 + q] - tv[q]; }
    P[0] = RAY_D[1] * e2[2] - RAY_D[2] * e2[1];
    P[1] = RAY_D[2] * e2[0] - RAY_D[0] * e2[2];
    P[2] = RAY_D[0] * e2[1] - RAY_D[1] * e2[0];
    float det = e1[0] * P[0] + e1[1] * P[1] + e1[2] * P[2];
    if ((double)det > -MT_EPS && (double)det < MT_EPS) return 0;
    m->inv_det = 1.f / det;
    return 1;
}

/* the u/v half for the column (ox, oy) */
static int mt_xy_col(const mt_tri *m, float ox, float oy, mt_col *s)
{
    const float *e1 = m->e1, *P = m->P;
    float inv_det = m->inv_det;
    float tx = ox - m->v1[0], ty = oy - m->v1[1];
    /* T.z * P.z is an exact signed zero: it cannot change a nonzero sum, and
     * a zero sum gives u = +-0 which passes/fails every compare identically. */
    float u = (tx * P[0] + ty * P[1] + 0.0f * P[2]) * inv_det;
    if (u < 0.f || u > 1.f) return 0;
    float qz = tx * e1[1] - ty * e1[0];
    float v = (RAY_D[0] * 0.0f + RAY_D[1] * 0.0f + RAY_D[2] * qz) * inv_det;
    if (v < 0.f || u + v > 1.f) return 0;
    for (int q = 0; q < 3; q++) { s->v1[q] = m->v1[q]; s->e1[q] = e1[q]; s->e2[q] = m->e2[q]; }
    s->inv_det = inv_det;
    s->tx = tx;
    s->ty = ty;
    return 1;
}

static int mt_xy(const float *tv, float ox, float oy, mt_col *s)
{
    mt_tri m;
    return mt_setup(tv, &m) && mt_xy_col(&m, ox, oy, s);
}

static int sgnf(float x) { return (x > 0.0f) - (x < 0.0f); }

/* Whether mt_t(s, oz) is monotone in oz over [oz_lo, oz_hi], and which way
 * (*dir <= 0: the test passes on a prefix of increasing oz, > 0: on a
 * suffix).  Rounding is monotone, so each float operation of mt_t is
 * monotone in a monotone operand: tz = oz - v1.z rises with oz; qx falls
 * with tz when e1.y > 0 (rises when < 0), qy rises with tz when e1.x > 0;
 * e2.x*qx and e2.y*qy are then monotone with directions
 * -sgn(e2.x)sgn(e1.y) and sgn(e2.y)sgn(e1.x), the left-to-right sum is
 * monotone when those two directions do not oppose (e2.z*qz is constant),
 * and the product by inv_det flips it when det < 0.  The magnitude guard
 * keeps every intermediate finite (no overflow, no NaN) over the range. */
static int mt_t_monotone(const mt_col *s, float oz_lo, float oz_hi, int *dir)
{
    const float lim = 1e15f;
    float tz0 = oz_lo - s->v1[2], tz1 = oz_hi - s->v1[2];
    if (!(fabsf(tz0) <= lim && fabsf(tz1) <= lim && fabsf(s->tx) <= lim && fabsf(s->ty) <= lim &&
          fabsf(s->inv_det) <= 1e7f))
        return 0;
    for (int q = 0; q < 3; q++)
        if (!(fabsf(s->e1[q]) <= lim && fabsf(s->e2[q]) <= lim)) return 0;
    int da = -sgnf(s->e2[0]) * sgnf(s->e1[1]);
    int db = sgnf(s->e2[1]) * sgnf(s->e1[0]);
    if (da * db < 0) return 0;
    *dir = (da ? da : db) * sgnf(s->inv_det);
    return 1;
}

static int mt_t(const mt_col *s, float oz)
{
    float tz = oz - s->v1[2];
    float qx = s->ty * s->e1[2] - tz * s->e1[1];
    float qy = tz * s->e1[0] - s->tx * s->e1[2];
    float qz = s->tx * s->e1[1] - s->ty * s->e1[0];
    float t = (s->e2[0] * qx + s->e2[1] * qy + s->e2[2] * qz) * s->inv_det;
    return (double)t > MT_EPS;
}

void orc_tri_box_batch(int64_t n, const float *c, const float *h, const float *tv, uint8_t *res)
{
    for (int64_t i = 0; i < n; i++) res[i] = (uint8_t)orc_tri_box_overlap(c + 3 * i, h + 3 * i, tv + 9 * i);
}

void orc_tri_ray_batch(int64_t n, const float *tv, const float *o, const float *d, uint8_t *hit,
                       float *t)
{
    for (int64_t i = 0; i < n; i++) {
        float out = 0.0f;
        hit[i] = (uint8_t)orc_tri_ray(tv + 9 * i, tv + 9 * i + 3, tv + 9 * i + 6, o + 3 * i, d + 3 * i, &out);
        t[i] = hit[i] ? out : 0.0f;
    }
}

void orc_tri_zray_split_batch(int64_t n, const float *tv, const float *o, uint8_t *hit)
{
    for (int64_t i = 0; i < n; i++) {
        mt_col s;
        hit[i] = (uint8_t)(mt_xy(tv + 9 * i, o[3 * i], o[3 * i + 1], &s) && mt_t(&s, o[3 * i + 2]));
    }
}

/* ------------------------------------------------------------------------ */
/* bbox and grid                                                             */
/* ------------------------------------------------------------------------ */

/* src/Object.cpp:441-450 (MinFloat3/MaxFloat3 includes/FloatVector.h:344-360)
 * and the pad at src/Object.cpp:572-583.  VectorMagnitude
 * (includes/FloatVector.h:323-326) sums in float and takes sqrt; the
 * float->double->float round trip of sqrt equals a float sqrt. */
void orc_bbox(const float *verts, int64_t nv, float bmin[3], float bmax[3], float *mms)
{
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int64_t i = 0; i < nv; i++) {
        const float *p = verts + 3 * i;
        for (int q = 0; q < 3; q++) {
            if (i == 0) { mn[q] = p[q]; mx[q] = p[q]; continue; }
            mn[q] = mn[q] < p[q] ? mn[q] : p[q];
            mx[q] = mx[q] > p[q] ? mx[q] : p[q];
        }
    }
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    float sq = d0 * d0 + d1 * d1 + d2 * d2;
    float mag = (float)sqrt((double)sq);
    double model_size = mag;
    float offset = (float)(0.001 * model_size);
    for (int q = 0; q < 3; q++) {
        bmax[q] = offset + mx[q];
        bmin[q] = mn[q] - offset;
    }
    float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
    float m12 = ey > ez ? ey : ez;
    *mms = ex > m12 ? ex : m12;
}

static int next_div4(int a) { return (a % 4 == 0) ? a : a + (4 - a % 4); }

/* src/Object.cpp:3094-3136 */
void orc_grid_setup(const float bmin[3], const float bmax[3], float mms,
                    int voxel_count, int voxel_count2, orc_grid *g)
{
    memset(g, 0, sizeof(*g));
    for (int q = 0; q < 3; q++) { g->bmin[q] = bmin[q]; g->bmax[q] = bmax[q]; }
    g->max_model_size = mms;
    float nominal = (float)((double)mms / (1.0 * voxel_count));
    for (int q = 0; q < 3; q++) {
        float extent = bmax[q] - bmin[q];
        int n = (int)(extent / nominal);
        if (n == 0) n++;
        n = next_div4(n);
        g->n[q] = n;
        g->g[q] = (float)((double)(bmax[q] - bmin[q]) / (n * 1.0));
        g->ext1[q] = (float)((double)g->g[q] / 2.0);
    }
    if (voxel_count2 > 0) {
        g->n2 = voxel_count2;
        for (int q = 0; q < 3; q++) {
            g->g2[q] = (float)((double)g->g[q] / (voxel_count2 * 1.0));
            g->ext2[q] = (float)((double)g->g2[q] / 2.0);
        }
    }
}

/* Level1 SAT centers: boxMidPoint = (p + 0.5)*boxExtents*2 + bBoxMin
 * (src/Object.cpp:2329-2331), evaluated in double, stored as float. */
void orc_l1_centers(const orc_grid *g, int a, float *out)
{
    for (int p = 0; p < g->n[a]; p++)
        out[p] = (float)((p + 0.5) * (double)g->ext1[a] * 2 + (double)g->bmin[a]);
}

/* Level1 ray origins: bBoxMin + (i + 0.5)*gridSize (src/Object.cpp:743-745). */
void orc_l1_ray_origins(const orc_grid *g, int a, float *out)
{
    for (int i = 0; i < g->n[a]; i++)
        out[i] = (float)((double)g->bmin[a] + (i + 0.5) * (double)g->g[a]);
}

/* The CUDA kernels' Level2 center and ray origin, one float expression for
 * both (cuda/CUDAClassifyTessellation.cu:423-425, 472-474):
 * (2p+1)*boxExtentsLevel2 + midPoint - boxExtentsLevel1, evaluated left to
 * right in float, midPoint = (float)((i + 0.5)*gridSize + bBoxMin) from the
 * host (src/Object.cpp:2566-2569), extents (float)(gridSize/2.0)
 * (src/Object.cpp:2553-2554).  No multiply-add contraction. */
float orc_l2_cuda_center(const orc_grid *g, int a, float mid, int p)
{
    float prod = (float)(2 * p + 1) * g->ext2[a];
    float sum = prod + mid;
    return sum - g->ext1[a];
}

/* Level2 SAT centers (src/Object.cpp:2369-2371, 2389-2391):
 * mid = (i + 0.5)*g + bBoxMin; center = (p + 0.5)*g2 + mid - ext1. */
void orc_l2_centers(const orc_grid *g, int a, float *out)
{
    int n2 = g->n2;
    for (int i = 0; i < g->n[a]; i++) {
        float mid = (float)((i + 0.5) * (double)g->g[a] + (double)g->bmin[a]);
        for (int p = 0; p < n2; p++)
            out[i * n2 + p] = g->cuda_l2 ? orc_l2_cuda_center(g, a, mid, p)
                                         : (float)((p + 0.5) * (double)g->g2[a] + (double)mid -
                                                   (double)g->ext1[a]);
    }
}

/* Level2 ray origins (src/Object.cpp:1134-1136, 1158-1160):
 * minPoint = i*g + bBoxMin in float; origin = minPoint + (p + 0.5)*g2. */
void orc_l2_ray_origins(const orc_grid *g, int a, float *out)
{
    int n2 = g->n2;
    for (int i = 0; i < g->n[a]; i++) {
        float minp = (float)i * g->g[a] + g->bmin[a];
        float mid = (float)((i + 0.5) * (double)g->g[a] + (double)g->bmin[a]);
        for (int p = 0; p < n2; p++)
            out[i * n2 + p] = g->cuda_l2 ? orc_l2_cuda_center(g, a, mid, p)
                                         : (float)((double)minp + (p + 0.5) * (double)g->g2[a]);
    }
}

/* src/Object.cpp:2272-2281: int((v - min)/(max - min) * nDiv), pulled back
 * by one when a vertex sits exactly on the max plane. */
int orc_vertex_index(float v, float mn, float mx, int n)
{
    float num = v - mn;
    float den = mx - mn;
    float r = num / den;
    float s = r * (float)n;
    int i = (int)s;
    if (i == n && v == mx) i--;
    return i;
}

/* Clipped index-AABB of one triangle (src/Object.cpp:2303-2318 and the loop
 * bounds :2320-2324).  Returns the number of voxels the reference tests. */
static int64_t tri_index_box(const float *tv, const orc_grid *g, int lo[3], int hi[3])
{
    int64_t vol = 1;
    for (int q = 0; q < 3; q++) {
        int i0 = orc_vertex_index(tv[q], g->bmin[q], g->bmax[q], g->n[q]);
        int i1 = orc_vertex_index(tv[3 + q], g->bmin[q], g->bmax[q], g->n[q]);
        int i2 = orc_vertex_index(tv[6 + q], g->bmin[q], g->bmax[q], g->n[q]);
        int mn = i0 < i1 ? i0 : i1; mn = mn < i2 ? mn : i2;
        int mx = i0 > i1 ? i0 : i1; mx = mx > i2 ? mx : i2;
        if (mn < 0) mn = 0;
        if (mx > g->n[q] - 1) mx = g->n[q] - 1;
        lo[q] = mn;
        hi[q] = mx;
        vol *= (mx >= mn) ? (int64_t)(mx - mn + 1) : 0;
    }
    return vol;
}

/* ------------------------------------------------------------------------ */
/* conservative column bounds for the all-triangle Level1 parity             */
/* ------------------------------------------------------------------------ */
/* For D = +z the MT acceptance (u >= 0, v >= 0, u + v <= 1 in float) is a
 * perturbed barycentric test.  With U = 2^-24 and |T| bounded over the grid,
 * the float u, v differ from the exact barycentric coordinates (u*, v*) of
 * the ray point w.r.t. the float edges e1, e2 by at most
 *    du = (E_un + E_det (1+2U)) / (|det| - E_det) + 3U      (same for v)
 * where E_det, E_un bound the rounding of det and of the u numerator.  An
 * accepted column therefore lies in {V1 + a e1 + b e2 : a >= -du, b >= -dv,
 * a + b <= 1 + 2U + du + dv}; the xy bbox of that triangle is returned,
 * widened by 'safety'.  Slivers with |det| near 1e-6 get a wide box, which
 * is exactly where float MT accepts rays well outside the triangle. */
int orc_ray_column_bounds(const float tv[9], const float bmin[3], const float bmax[3],
                          double safety, double lo[2], double hi[2])
{
    float e1[3], e2[3], P[3];
    for (int q = 0; q < 3; q++) { e1[q] = tv[3 + q] - tv[q]; e2[q] = tv[6 + q] - tv[q]; }
    P[0] = RAY_D[1] * e2[2] - RAY_D[2] * e2[1];
    P[1] = RAY_D[2] * e2[0] - RAY_D[0] * e2[2];
    P[2] = RAY_D[0] * e2[1] - RAY_D[1] * e2[0];
    float det = e1[0] * P[0] + e1[1] * P[1] + e1[2] * P[2];
    if ((double)det > -MT_EPS && (double)det < MT_EPS) return 0;
    const double U = 5.9604644775390625e-08;
    double ad = fabs((double)det);
    double tmx = fmax(fabs((double)bmin[0] - tv[0]), fabs((double)bmax[0] - tv[0]));
    double tmy = fmax(fabs((double)bmin[1] - tv[1]), fabs((double)bmax[1] - tv[1]));
    double e1x = e1[0], e1y = e1[1], e2x = e2[0], e2y = e2[1];
    double edet = U * (fabs(e1x * e2y) + fabs(e1y * e2x) + ad);
    double eun = U * (3.0 * (fabs(e2y) * tmx + fabs(e2x) * tmy) + 2.0 * ad);
    double evn = U * (3.0 * (fabs(e1y) * tmx + fabs(e1x) * tmy) + 2.0 * ad);
    double den = ad - edet;
    if (den <= 0.5 * ad) { /* det not even sign-reliable: whole grid */
        lo[0] = bmin[0]; lo[1] = bmin[1];
        hi[0] = bmax[0]; hi[1] = bmax[1];
        return 1;
    }
    double du = safety * ((eun + edet * (1 + 2 * U)) / den + 3 * U);
    double dv = safety * ((evn + edet * (1 + 2 * U)) / den + 3 * U);
    double ds = du + dv + safety * 2 * U;
    double ca[3] = {-du, 1 + ds + dv, -du};
    double cb[3] = {-dv, -dv, 1 + ds + du};
    for (int ax = 0; ax < 2; ax++) {
        double ea = ax ? e1y : e1x, eb = ax ? e2y : e2x, base = tv[ax];
        double mn = 1e300, mx = -1e300;
        for (int c = 0; c < 3; c++) {
            double x = base + ca[c] * ea + cb[c] * eb;
            mn = x < mn ? x : mn;
            mx = x > mx ? x : mx;
        }
        double slop = 1e-12 * (fabs(mn) + fabs(mx) + 1.0);
        lo[ax] = mn - slop;
        hi[ax] = mx + slop;
    }
    return 1;
}

/* first index with tab[i] >= x (n if none) */
static int lower_idx(const float *tab, int n, double x)
{
    int a = 0, b = n;
    while (a < b) {
        int m = (a + b) / 2;
        if ((double)tab[m] < x) a = m + 1; else b = m;
    }
    return a;
}

/* ------------------------------------------------------------------------ */
/* Level1 GL-stencil-equivalent fill (src/Object.cpp:1446-1601): the rule of  */
/* orc_stencil_counts in gpv_oracle.h                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t X[3], Y[3];               /* snapped window coordinates, counter-clockwise */
    int64_t pxlo, pxhi, pylo, pyhi;   /* samples whose centers lie in the snapped bbox */
    double v0[3];                     /* first vertex */
    double nx, ny, nz;                /* plane normal (e1 x e2 of the float vertices) */
} st_tri;

#define ST_MAX_SAMPLES (1 << 20)

static int64_t floor_div(int64_t a, int64_t b) /* b > 0 */
{
    int64_t q = a / b;
    return (a % b != 0 && a < 0) ? q - 1 : q;
}

/* window coordinate of a vertex coordinate, in 1/256 sample units */
static int64_t st_snap(float x, float lo, float hi, int samples)
{
    double w = ((double)x - (double)lo) / ((double)hi - (double)lo) * (double)samples;
    /* vertices far outside the bbox: bounded so that edge functions stay in 64 bits */
    w = w < -4194304.0 ? -4194304.0 : (w > 4194304.0 ? 4194304.0 : w);
    return (int64_t)rint(w * 256.0);
}

/* 0 when the triangle covers no sample in any slice */
static int st_prepare(const float *tv, const orc_grid *g, st_tri *s)
{
    const int W = 2 * g->n[0], H = 2 * g->n[1];
    for (int i = 0; i < 3; i++) {
        s->X[i] = st_snap(tv[3 * i], g->bmin[0], g->bmax[0], W);
        s->Y[i] = st_snap(tv[3 * i + 1], g->bmin[1], g->bmax[1], H);
    }
    int64_t A = (s->X[1] - s->X[0]) * (s->Y[2] - s->Y[0]) - (s->Y[1] - s->Y[0]) * (s->X[2] - s->X[0]);
    if (A == 0) return 0;
    if (A < 0) {
        int64_t t = s->X[1]; s->X[1] = s->X[2]; s->X[2] = t;
        t = s->Y[1]; s->Y[1] = s->Y[2]; s->Y[2] = t;
    }
    double e1[3], e2[3];
    for (int q = 0; q < 3; q++) {
        s->v0[q] = (double)tv[q];
        e1[q] = (double)tv[3 + q] - (double)tv[q];
        e2[q] = (double)tv[6 + q] - (double)tv[q];
    }
    s->nx = e1[1] * e2[2] - e1[2] * e2[1];
    s->ny = e1[2] * e2[0] - e1[0] * e2[2];
    s->nz = e1[0] * e2[1] - e1[1] * e2[0];
    if (s->nz == 0.0) return 0;
    int64_t mnx = s->X[0], mxx = s->X[0], mny = s->Y[0], mxy = s->Y[0];
    for (int i = 1; i < 3; i++) {
        if (s->X[i] < mnx) mnx = s->X[i];
        if (s->X[i] > mxx) mxx = s->X[i];
        if (s->Y[i] < mny) mny = s->Y[i];
        if (s->Y[i] > mxy) mxy = s->Y[i];
    }
    s->pxlo = floor_div(mnx - 128 + 255, 256);
    s->pxhi = floor_div(mxx - 128, 256);
    s->pylo = floor_div(mny - 128 + 255, 256);
    s->pyhi = floor_div(mxy - 128, 256);
    if (s->pxlo < 0) s->pxlo = 0;
    if (s->pylo < 0) s->pylo = 0;
    if (s->pxhi > W - 1) s->pxhi = W - 1;
    if (s->pyhi > H - 1) s->pyhi = H - 1;
    return s->pxlo <= s->pxhi && s->pylo <= s->pyhi;
}

/* edge functions of the snapped triangle at the sample, top-left rule */
static int st_covers(const st_tri *s, int64_t px, int64_t py)
{
    const int64_t sx = 256 * px + 128, sy = 256 * py + 128;
    for (int e = 0; e < 3; e++) {
        const int a = e, b = e == 2 ? 0 : e + 1;
        const int64_t dx = s->X[b] - s->X[a], dy = s->Y[b] - s->Y[a];
        const int64_t E = dx * (sy - s->Y[a]) - dy * (sx - s->X[a]);
        if (E < 0) return 0;
        if (E == 0 && !(dy < 0 || (dy == 0 && dx < 0))) return 0;
    }
    return 1;
}

static int lower_idx(const float *tab, int n, double x);

/* first slice whose clip plane keeps the triangle's point over the sample
 * (z_T <= zs); nz when none does */
static int st_slice(const st_tri *s, const orc_grid *g, const float *rz, int64_t px, int64_t py)
{
    const double l = g->bmin[0], r = g->bmax[0], b = g->bmin[1], t = g->bmax[1];
    const double x = l + ((double)px + 0.5) * ((r - l) / (double)(2 * g->n[0]));
    const double y = b + ((double)py + 0.5) * ((t - b) / (double)(2 * g->n[1]));
    const double zt = s->v0[2] - (s->nx * (x - s->v0[0]) + s->ny * (y - s->v0[1])) / s->nz;
    return lower_idx(rz, g->n[2], zt);
}

int orc_stencil_counts(const float *tris, int64_t T, const orc_grid *g, int slice, int32_t *counts)
{
    const int64_t W = 2 * (int64_t)g->n[0], H = 2 * (int64_t)g->n[1];
    if (W > ST_MAX_SAMPLES || H > ST_MAX_SAMPLES) return -1;
    memset(counts, 0, sizeof(int32_t) * (size_t)(W * H));
    float *rz = (float *)xcalloc((size_t)g->n[2], sizeof(float));
    orc_l1_ray_origins(g, 2, rz);
    for (int64_t t = 0; t < T; t++) {
        st_tri s;
        if (!st_prepare(tris + 9 * t, g, &s)) continue;
        for (int64_t py = s.pylo; py <= s.pyhi; py++)
            for (int64_t px = s.pxlo; px <= s.pxhi; px++)
                if (st_covers(&s, px, py) && st_slice(&s, g, rz, px, py) <= slice) counts[py * W + px]++;
    }
    free(rz);
    return 0;
}

/* every covered sample toggles its column from its first slice up; the
 * running parity per column is then the stencil's parity per slice, and
 * the 2 x 2 samples of a voxel vote (src/Object.cpp:1580-1601) */
int orc_stencil_inside(const float *tris, int64_t T, const orc_grid *g, uint8_t *inside)
{
    const int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2], W = 2 * nx, H = 2 * ny;
    if (W > ST_MAX_SAMPLES || H > ST_MAX_SAMPLES) return -1;
    float *rz = (float *)xcalloc((size_t)nz, sizeof(float));
    orc_l1_ray_origins(g, 2, rz);
    uint8_t *tog = (uint8_t *)xcalloc((size_t)(nz * H * W), 1);
    for (int64_t t = 0; t < T; t++) {
        st_tri s;
        if (!st_prepare(tris + 9 * t, g, &s)) continue;
        for (int64_t py = s.pylo; py <= s.pyhi; py++)
            for (int64_t px = s.pxlo; px <= s.pxhi; px++) {
                if (!st_covers(&s, px, py)) continue;
                const int k = st_slice(&s, g, rz, px, py);
                if (k < nz) tog[((int64_t)k * H + py) * W + px] ^= 1;
            }
    }
    for (int64_t q = 0; q < H * W; q++) {
        uint8_t p = 0;
        for (int64_t k = 0; k < nz; k++) {
            p ^= tog[k * H * W + q];
            tog[k * H * W + q] = p;
        }
    }
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                const uint8_t *a = tog + (k * H + 2 * j) * W + 2 * i;
                const int odd = a[0] + a[1] + a[W] + a[W + 1];
                inside[(k * ny + j) * nx + i] = (uint8_t)(odd > 2);   /* > numSamples^2 / 2.0 */
            }
    free(tog);
    free(rz);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* pipeline                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    const float *tris;
    int64_t T;
    const orc_grid *g;
    int nt;
    int64_t nx, ny, nz, n1;
    float *cx, *cy, *cz;     /* L1 SAT centers */
    float *rx, *ry, *rz;     /* L1 ray origins */
    int mode;                /* ORC_EXACT: literal loops; ORC_FAST, ORC_PLAN: result-identical forms */
} ctx_t;

/* Level1 ray parity over ALL triangles (src/Object.cpp:716-779), literal. */
static void l1_parity_exact(const ctx_t *c, uint8_t *par)
{
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(c->nt)
    for (int64_t kj = 0; kj < nz * ny; kj++) {
        int64_t k = kj / ny, j = kj % ny;
        for (int64_t i = 0; i < nx; i++) {
            float o[3] = {c->rx[i], c->ry[j], c->rz[k]};
            int hits = 0;
            for (int64_t t = 0; t < c->T; t++) {
                const float *tv = c->tris + 9 * t;
                float tp;
                if (orc_tri_ray(tv, tv + 3, tv + 6, o, RAY_D, &tp)) hits++;
            }
            par[k * ny * nx + j * nx + i] = (uint8_t)(hits % 2 == 1);
        }
    }
}

/* Level1 ray parity, column factorised: triangles are binned to the columns
 * their conservative bound admits, det/u/v run once per (triangle, column)
 * and only t runs per voxel. */
static void l1_parity_fast(const ctx_t *c, uint8_t *par)
{
    int64_t nx = c->nx, ny = c->ny, nz = c->nz, ncol = nx * ny;
    int nt = c->nt;
    vec64 *loc = (vec64 *)xcalloc((size_t)nt, sizeof(vec64));
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
#pragma omp for schedule(static)
        for (int64_t t = 0; t < c->T; t++) {
            const float *tv = c->tris + 9 * t;
            double lo[2], hi[2];
            if (!orc_ray_column_bounds(tv, c->g->bmin, c->g->bmax, 8.0, lo, hi)) continue;
            int ia = lower_idx(c->rx, (int)nx, lo[0]);
            int ib = lower_idx(c->rx, (int)nx, hi[0]);  /* first > hi is <= this */
            while (ib < nx && (double)c->rx[ib] <= hi[0]) ib++;
            int ja = lower_idx(c->ry, (int)ny, lo[1]);
            int jb = lower_idx(c->ry, (int)ny, hi[1]);
            while (jb < ny && (double)c->ry[jb] <= hi[1]) jb++;
            for (int j = ja; j < jb; j++)
                for (int i = ia; i < ib; i++)
                    v64_push(&loc[tid], ((int64_t)(j * nx + i) << 32) | t);
        }
    }
    int64_t *cnt = (int64_t *)xcalloc((size_t)ncol + 1, sizeof(int64_t));
    int64_t total = 0;
    for (int th = 0; th < nt; th++) {
        for (int64_t q = 0; q < loc[th].n; q++) cnt[(loc[th].a[q] >> 32) + 1]++;
        total += loc[th].n;
    }
    for (int64_t q = 0; q < ncol; q++) cnt[q + 1] += cnt[q];
    int32_t *lst = (int32_t *)xcalloc((size_t)total, sizeof(int32_t));
    int64_t *fill = (int64_t *)xcalloc((size_t)ncol, sizeof(int64_t));
    for (int th = 0; th < nt; th++) {
        for (int64_t q = 0; q < loc[th].n; q++) {
            int64_t col = loc[th].a[q] >> 32;
            lst[cnt[col] + fill[col]++] = (int32_t)(loc[th].a[q] & 0xffffffff);
        }
        free(loc[th].a);
    }
    free(loc);
    free(fill);
#pragma omp parallel num_threads(nt)
    {
        mt_col *hits = NULL;
        int64_t cap = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t col = 0; col < ncol; col++) {
            int64_t i = col % nx, j = col / nx;
            int64_t n = cnt[col + 1] - cnt[col];
            if (n > cap) { cap = n; hits = (mt_col *)realloc(hits, (size_t)cap * sizeof(mt_col)); }
            int64_t nh = 0;
            for (int64_t q = 0; q < n; q++) {
                const float *tv = c->tris + 9 * (int64_t)lst[cnt[col] + q];
                if (mt_xy(tv, c->rx[i], c->ry[j], &hits[nh])) nh++;
            }
            for (int64_t k = 0; k < nz; k++) {
                int p = 0;
                for (int64_t h = 0; h < nh; h++) p ^= mt_t(&hits[h], c->rz[k]);
                par[k * ny * nx + j * nx + i] = (uint8_t)p;
            }
        }
        free(hits);
    }
    free(cnt);
    free(lst);
}

/* sat_sweep callback of the Level1 sweep: one bin entry per overlap */
typedef struct { vec64 *out; int p0, q0, s0; int64_t nx, ny, t; } l1_emit_t;

static void l1_emit(void *ud, int p, int q, int s)
{
    l1_emit_t *e = (l1_emit_t *)ud;
    int64_t v = (int64_t)(e->s0 + s) * e->ny * e->nx + (int64_t)(e->q0 + q) * e->nx + (e->p0 + p);
    v64_push(e->out, (v << 32) | e->t);
}

/* Level1 SAT sweep + bins (src/Object.cpp:2256-2342 with AddPointToVoxel
 * :2705-2748).  Bins list triangle ids in ascending order, as the
 * reference's triangle-outer loop appends them. */
static void l1_sat(const ctx_t *c, orc_result *r)
{
    int nt = c->nt;
    int64_t nx = c->nx, ny = c->ny;
    vec64 *loc = (vec64 *)xcalloc((size_t)nt, sizeof(vec64));
    int64_t tests = 0;
    const float *h = c->g->ext1;
#pragma omp parallel num_threads(nt) reduction(+ : tests)
    {
        int tid = omp_get_thread_num();
        int nth = omp_get_num_threads();
        int64_t t0 = c->T * tid / nth, t1 = c->T * (tid + 1) / nth;
        bytes_t tmp = {0};
        l1_emit_t em = {&loc[tid], 0, 0, 0, nx, ny, 0};
        for (int64_t t = t0; t < t1; t++) {
            const float *tv = c->tris + 9 * t;
            int lo[3], hi[3];
            tests += tri_index_box(tv, c->g, lo, hi);
            if (c->mode == ORC_FAST) {
                em.p0 = lo[0]; em.q0 = lo[1]; em.s0 = lo[2]; em.t = t;
                sat_sweep(tv, h, c->cx + lo[0], hi[0] - lo[0] + 1, c->cy + lo[1], hi[1] - lo[1] + 1,
                          c->cz + lo[2], hi[2] - lo[2] + 1, &tmp, l1_emit, &em);
                continue;
            }
            for (int p = lo[0]; p <= hi[0]; p++)
                for (int q = lo[1]; q <= hi[1]; q++)
                    for (int s = lo[2]; s <= hi[2]; s++) {
                        float ctr[3] = {c->cx[p], c->cy[q], c->cz[s]};
                        if (orc_tri_box_overlap(ctr, h, tv))
                            v64_push(&loc[tid], ((int64_t)(s * ny * nx + q * nx + p) << 32) | t);
                    }
        }
        free(tmp.a);
    }
    r->tests_l1 = tests;
    int64_t n1 = c->n1;
    int64_t *off = (int64_t *)xcalloc((size_t)n1 + 1, sizeof(int64_t));
    int64_t total = 0;
    for (int th = 0; th < nt; th++) {
        for (int64_t q = 0; q < loc[th].n; q++) off[(loc[th].a[q] >> 32) + 1]++;
        total += loc[th].n;
    }
    int maxc = 0;
    for (int64_t v = 0; v < n1; v++) {
        if (off[v + 1] > maxc) maxc = (int)off[v + 1];
        off[v + 1] += off[v];
    }
    int32_t *lst = (int32_t *)xcalloc((size_t)total, sizeof(int32_t));
    int64_t *fill = (int64_t *)xcalloc((size_t)n1, sizeof(int64_t));
    for (int th = 0; th < nt; th++) {  /* threads in order keeps ids ascending */
        for (int64_t q = 0; q < loc[th].n; q++) {
            int64_t v = loc[th].a[q] >> 32;
            lst[off[v] + fill[v]++] = (int32_t)(loc[th].a[q] & 0xffffffff);
        }
        free(loc[th].a);
    }
    free(loc);
    free(fill);
    r->list_off = off;
    r->list_tri = lst;
    r->pairs = total;
    r->max_tris_per_voxel = maxc;
}

/* Level2 ray parity, literal (src/Object.cpp:1111-1207): candidates are the
 * bins of the column at z' >= k1, each triangle counted once. */
static void l2_parity_exact(const ctx_t *c, const orc_result *r, const float *ox,
                            const float *oy, const float *oz, uint8_t *par)
{
    int n2 = c->g->n2;
    int64_t n23 = (int64_t)n2 * n2 * n2, nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel num_threads(c->nt)
    {
        uint8_t *seen = (uint8_t *)xcalloc((size_t)c->T, 1);
        vec64 touched = {0};
#pragma omp for schedule(dynamic, 16)
        for (int64_t b = 0; b < r->nb; b++) {
            int64_t l1 = r->bidx[b];
            int64_t k1 = l1 / (nx * ny), ij = l1 - k1 * nx * ny, j1 = ij / nx, i1 = ij % nx;
            for (int s = 0; s < n2; s++)
                for (int q = 0; q < n2; q++)
                    for (int p = 0; p < n2; p++) {
                        float o[3] = {ox[i1 * n2 + p], oy[j1 * n2 + q], oz[k1 * n2 + s]};
                        int hits = 0;
                        /* the CUDA kernel's candidates: every bin of the column */
                        for (int64_t z = c->g->cuda_l2 ? 0 : k1; z < nz; z++) {
                            int64_t v = z * nx * ny + ij;
                            for (int64_t e = r->list_off[v]; e < r->list_off[v + 1]; e++) {
                                int32_t t = r->list_tri[e];
                                if (seen[t]) continue;
                                const float *tv = c->tris + 9 * (int64_t)t;
                                float tp;
                                if (orc_tri_ray(tv, tv + 3, tv + 6, o, RAY_D, &tp)) {
                                    hits++;
                                    seen[t] = 1;
                                    v64_push(&touched, t);
                                }
                            }
                        }
                        for (int64_t e = 0; e < touched.n; e++) seen[touched.a[e]] = 0;
                        touched.n = 0;
                        par[b * n23 + ((int64_t)s * n2 + q) * n2 + p] = (uint8_t)(hits % 2 == 1);
                    }
        }
        free(seen);
        free(touched.a);
    }
}

/* Level2 ray parity, column factorised.  For column (i1, j1) let zmax(t) be
 * the highest bin holding t; t is a candidate of boundary voxel k1 iff
 * zmax(t) >= k1, which is the deduplicated union of the literal loop. */
static void l2_parity_fast(const ctx_t *c, const orc_result *r, const float *ox,
                           const float *oy, const float *oz, uint8_t *par)
{
    int n2 = c->g->n2;
    int64_t n23 = (int64_t)n2 * n2 * n2, nx = c->nx, ny = c->ny, nz = c->nz;
    /* group boundary voxels by column: consecutive b with equal column after
     * a sort by (column, z) */
    int64_t nb = r->nb;
    int64_t *order = (int64_t *)xcalloc((size_t)nb, sizeof(int64_t));
    {
        int64_t ncol = nx * ny;
        int64_t *cc = (int64_t *)xcalloc((size_t)ncol + 1, sizeof(int64_t));
        for (int64_t b = 0; b < nb; b++) cc[r->bidx[b] % ncol + 1]++;
        for (int64_t q = 0; q < ncol; q++) cc[q + 1] += cc[q];
        for (int64_t b = 0; b < nb; b++) order[cc[r->bidx[b] % ncol]++] = b;  /* z ascending */
        free(cc);
    }
    int64_t *starts = (int64_t *)xcalloc((size_t)nb + 1, sizeof(int64_t));
    int64_t ng = 0;
    for (int64_t q = 0; q < nb; q++)
        if (q == 0 || r->bidx[order[q]] % (nx * ny) != r->bidx[order[q - 1]] % (nx * ny))
            starts[ng++] = q;
    starts[ng] = nb;
    /* ray origins along z non-decreasing: then a column's layers, boundary
     * voxels by ascending z and their sub-layers, have non-decreasing oz */
    int zmono = 1;
    for (int64_t k = 1; k < nz * n2; k++) zmono &= oz[k - 1] <= oz[k];
#pragma omp parallel num_threads(c->nt)
    {
        int32_t *zmax = (int32_t *)xcalloc((size_t)c->T, sizeof(int32_t));
        for (int64_t t = 0; t < c->T; t++) zmax[t] = -1;
        vec64 uniq = {0};
        mt_col *hs = NULL;
        int32_t *hz = NULL;
        mt_tri *ut = NULL;      /* set-up of the column's candidates that pass the det test */
        int32_t *uz = NULL;
        int64_t hcap = 0;
        float *ozl = NULL;      /* the group's layers: l = (m - m0)*n2 + s */
        int64_t *kk = NULL;     /* k1 of the group's boundary voxels, ascending */
        uint8_t *flip = NULL;   /* parity toggles at layer boundaries */
        int64_t lcap = 0;
#pragma omp for schedule(dynamic, 4)
        for (int64_t gi = 0; gi < ng; gi++) {
            int64_t ij = r->bidx[order[starts[gi]]] % (nx * ny);
            int64_t i1 = ij % nx, j1 = ij / nx;
            int64_t m0 = starts[gi], nbg = starts[gi + 1] - m0, L = nbg * n2;
            if (L + 1 > lcap) {
                lcap = L + 1;
                ozl = (float *)realloc(ozl, (size_t)lcap * sizeof(float));
                kk = (int64_t *)realloc(kk, (size_t)lcap * sizeof(int64_t));
                flip = (uint8_t *)realloc(flip, (size_t)lcap);
            }
            for (int64_t m = 0; m < nbg; m++) {
                kk[m] = r->bidx[order[m0 + m]] / (nx * ny);
                for (int s = 0; s < n2; s++) ozl[m * n2 + s] = oz[kk[m] * n2 + s];
            }
            for (int64_t z = 0; z < nz; z++) {
                int64_t v = z * nx * ny + ij;
                for (int64_t e = r->list_off[v]; e < r->list_off[v + 1]; e++) {
                    int32_t t = r->list_tri[e];
                    if (zmax[t] < 0) v64_push(&uniq, t);
                    zmax[t] = (int32_t)z;
                }
            }
            if (uniq.n > hcap) {
                hcap = uniq.n;
                hs = (mt_col *)realloc(hs, (size_t)hcap * sizeof(mt_col));
                hz = (int32_t *)realloc(hz, (size_t)hcap * sizeof(int32_t));
                ut = (mt_tri *)realloc(ut, (size_t)hcap * sizeof(mt_tri));
                uz = (int32_t *)realloc(uz, (size_t)hcap * sizeof(int32_t));
            }
            int64_t nu = 0;
            for (int64_t e = 0; e < uniq.n; e++) {
                int32_t t = (int32_t)uniq.a[e];
                if (mt_setup(c->tris + 9 * (int64_t)t, &ut[nu])) uz[nu++] = zmax[t];
            }
            for (int q = 0; q < n2; q++)
                for (int p = 0; p < n2; p++) {
                    int64_t nh = 0;
                    for (int64_t e = 0; e < nu; e++)
                        if (mt_xy_col(&ut[e], ox[i1 * n2 + p], oy[j1 * n2 + q], &hs[nh])) hz[nh++] = uz[e];
                    /* parity at layer l = XOR over candidates h (hz[h] >= k1 of
                     * l's voxel) of mt_t(h, ozl[l]); each candidate's passing
                     * layers are one run when mt_t is monotone in oz (a
                     * binary search finds it), else they are tested one by
                     * one; runs are toggled at their ends and summed by a
                     * prefix XOR */
                    memset(flip, 0, (size_t)L + 1);
                    for (int64_t h = 0; h < nh; h++) {
                        int64_t lend = L;
                        if (!c->g->cuda_l2) {   /* voxels with k1 <= hz[h]: a prefix */
                            int64_t a = 0, bb = nbg;
                            while (a < bb) {
                                int64_t mid = (a + bb) / 2;
                                if (kk[mid] <= hz[h]) a = mid + 1; else bb = mid;
                            }
                            lend = a * n2;
                        }
                        if (lend == 0) continue;
                        int dir = 0;
                        if (c->mode == ORC_FAST && zmono && mt_t_monotone(&hs[h], ozl[0], ozl[lend - 1], &dir)) {
                            int64_t a = 0, bb = lend;
                            if (dir <= 0) {   /* passes on [0, first failure) */
                                while (a < bb) {
                                    int64_t mid = (a + bb) / 2;
                                    if (mt_t(&hs[h], ozl[mid])) a = mid + 1; else bb = mid;
                                }
                                flip[0] ^= 1;
                                flip[a] ^= 1;
                            } else {          /* passes on [first pass, lend) */
                                while (a < bb) {
                                    int64_t mid = (a + bb) / 2;
                                    if (mt_t(&hs[h], ozl[mid])) bb = mid; else a = mid + 1;
                                }
                                flip[a] ^= 1;
                                flip[lend] ^= 1;
                            }
                        } else {
                            for (int64_t l = 0; l < lend; l++)
                                if (mt_t(&hs[h], ozl[l])) { flip[l] ^= 1; flip[l + 1] ^= 1; }
                        }
                    }
                    int pp = 0;
                    for (int64_t m = 0; m < nbg; m++) {
                        int64_t b = order[m0 + m];
                        for (int s = 0; s < n2; s++) {
                            pp ^= flip[m * n2 + s];
                            par[b * n23 + ((int64_t)s * n2 + q) * n2 + p] = (uint8_t)pp;
                        }
                    }
                }
            for (int64_t e = 0; e < uniq.n; e++) zmax[uniq.a[e]] = -1;
            uniq.n = 0;
        }
        free(zmax);
        free(uniq.a);
        free(hs);
        free(hz);
        free(ut);
        free(uz);
        free(ozl);
        free(kk);
        free(flip);
    }
    free(order);
    free(starts);
}

/* sat_sweep callback of the Level2 sweep: state 2 for the sub-voxel */
typedef struct { uint8_t *blk; int n2; } l2_emit_t;

static void l2_emit(void *ud, int p, int q, int s)
{
    l2_emit_t *e = (l2_emit_t *)ud;
    e->blk[((int64_t)s * e->n2 + q) * e->n2 + p] = 2;
}

/* Level2 SAT (src/Object.cpp:2344-2424): every triangle of the parent's bin
 * against every Level2 voxel; a hit sets state 2 in l2 (over the parity). */
static int64_t l2_sat(const ctx_t *c, const orc_result *r, const float *sx, const float *sy,
                      const float *sz, uint8_t *l2)
{
    int n2 = c->g->n2;
    int64_t n23 = (int64_t)n2 * n2 * n2, nx = c->nx, ny = c->ny;
    int64_t tests = 0;
    const float *h = c->g->ext2;
#pragma omp parallel num_threads(c->nt) reduction(+ : tests)
    {
        bytes_t tmp = {0};
#pragma omp for schedule(dynamic, 16)
        for (int64_t b = 0; b < r->nb; b++) {
            int64_t l1 = r->bidx[b];
            int64_t k1 = l1 / (nx * ny), ij = l1 - k1 * nx * ny, j1 = ij / nx, i1 = ij % nx;
            tests += (r->list_off[l1 + 1] - r->list_off[l1]) * n23;
            l2_emit_t em = {l2 + b * n23, n2};
            for (int64_t e = r->list_off[l1]; e < r->list_off[l1 + 1]; e++) {
                const float *tv = c->tris + 9 * (int64_t)r->list_tri[e];
                if (c->mode == ORC_FAST) {
                    sat_sweep(tv, h, sx + i1 * n2, n2, sy + j1 * n2, n2, sz + k1 * n2, n2, &tmp, l2_emit, &em);
                    continue;
                }
                for (int p = 0; p < n2; p++)
                    for (int q = 0; q < n2; q++)
                        for (int s = 0; s < n2; s++) {
                            float ctr[3] = {sx[i1 * n2 + p], sy[j1 * n2 + q], sz[k1 * n2 + s]};
                            if (orc_tri_box_overlap(ctr, h, tv))
                                l2[b * n23 + ((int64_t)s * n2 + q) * n2 + p] = 2;
                        }
            }
        }
        free(tmp.a);
    }
    return tests;
}

int orc_voxelize(const float *tris, int64_t T, const orc_grid *g, int mode, int threads,
                 orc_result *r)
{
    memset(r, 0, sizeof(*r));
    ctx_t c;
    c.tris = tris;
    c.T = T;
    c.g = g;
    c.nt = threads > 0 ? threads : omp_get_max_threads();
    c.mode = mode;
    c.nx = g->n[0]; c.ny = g->n[1]; c.nz = g->n[2];
    c.n1 = c.nx * c.ny * c.nz;
    float *tab[6];
    for (int q = 0; q < 3; q++) {
        tab[q] = (float *)xcalloc((size_t)g->n[q], sizeof(float));
        tab[3 + q] = (float *)xcalloc((size_t)g->n[q], sizeof(float));
        orc_l1_centers(g, q, tab[q]);
        orc_l1_ray_origins(g, q, tab[3 + q]);
    }
    c.cx = tab[0]; c.cy = tab[1]; c.cz = tab[2];
    c.rx = tab[3]; c.ry = tab[4]; c.rz = tab[5];
    r->n1 = c.n1;
    r->n2 = g->n2;

    double t0 = omp_get_wtime();
    uint8_t *par = (uint8_t *)xcalloc((size_t)c.n1, 1);
    if (g->l1_stencil) {
        if (orc_stencil_inside(tris, T, g, par) != 0) {
            free(par);
            for (int q = 0; q < 6; q++) free(tab[q]);
            return -2;
        }
    } else if (mode == ORC_EXACT) {
        l1_parity_exact(&c, par);
    } else {
        l1_parity_fast(&c, par);
    }
    double t1 = omp_get_wtime();
    l1_sat(&c, r);
    double t2 = omp_get_wtime();

    r->l1 = (uint8_t *)xcalloc((size_t)c.n1, 1);
    r->prefix = (int32_t *)xcalloc((size_t)c.n1, sizeof(int32_t));
    int64_t nb = 0;
    for (int64_t v = 0; v < c.n1; v++) {
        uint8_t s = (r->list_off[v + 1] > r->list_off[v]) ? 2 : par[v];
        r->l1[v] = s;
        r->prefix[v] = (int32_t)nb;   /* src/Object.cpp:3270-3279 */
        if (s == 2) nb++;
        if (s == 1) r->l1_inside++;
    }
    free(par);
    r->nb = nb;
    r->l1_boundary = nb;
    r->bidx = (int32_t *)xcalloc((size_t)nb, sizeof(int32_t));
    nb = 0;
    for (int64_t v = 0; v < c.n1; v++)
        if (r->l1[v] == 2) r->bidx[nb++] = (int32_t)v;
    double t3 = omp_get_wtime();
    double t4 = t3, t5 = t3;

    if (g->n2 > 0) {
        int n2 = g->n2;
        int64_t n23 = (int64_t)n2 * n2 * n2;
        float *l2t[6];
        for (int q = 0; q < 3; q++) {
            l2t[q] = (float *)xcalloc((size_t)g->n[q] * n2, sizeof(float));
            l2t[3 + q] = (float *)xcalloc((size_t)g->n[q] * n2, sizeof(float));
            orc_l2_centers(g, q, l2t[q]);
            orc_l2_ray_origins(g, q, l2t[3 + q]);
        }
        /* the parity writes 0/1 into every Level2 voxel, then the SAT
         * overwrites its hits with 2: state = SAT hit ? 2 : parity */
        r->l2 = (uint8_t *)xcalloc((size_t)(r->nb * n23), 1);
        if (mode == ORC_EXACT) l2_parity_exact(&c, r, l2t[3], l2t[4], l2t[5], r->l2);
        else l2_parity_fast(&c, r, l2t[3], l2t[4], l2t[5], r->l2);
        t4 = omp_get_wtime();
        r->tests_l2 = l2_sat(&c, r, l2t[0], l2t[1], l2t[2], r->l2);
        t5 = omp_get_wtime();
        int64_t ins2 = 0, bnd2 = 0;
        const int64_t nl2 = r->nb * n23;
        const uint8_t *l2s = r->l2;
#pragma omp parallel for schedule(static) num_threads(c.nt) reduction(+ : ins2, bnd2)
        for (int64_t q = 0; q < nl2; q++) {
            ins2 += l2s[q] == 1;     /* int(v) % 2 == 1, src/Object.cpp:3366 */
            bnd2 += l2s[q] == 2;
        }
        r->l2_inside = ins2;
        r->l2_boundary = bnd2;
        for (int q = 0; q < 6; q++) free(l2t[q]);
    }
    r->t_stage[0] = t1 - t0;
    r->t_stage[1] = t2 - t1;
    r->t_stage[2] = t3 - t2;
    r->t_stage[3] = t4 - t3;
    r->t_stage[4] = t5 - t4;
    for (int q = 0; q < 6; q++) free(tab[q]);
    return 0;
}

/* ---------------------------------------------------------------------------
 * normals (the part of the path that feeds Level1Normal.raw / Level2Normal.raw)
 * ------------------------------------------------------------------------- */

/* VectorNormalize, includes/FloatVector.h:332-341: float magnitude of the
 * left-to-right float sum of squares; components divided when nonzero. */
void orc_normalize(float v[3])
{
    float m = (float)sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    if (m != 0) {
        v[0] /= m;
        v[1] /= m;
        v[2] /= m;
    }
}

/* Face normal of the flat soup triangle, src/Object.cpp:3233-3240 (and the
 * load-time faceNormal of :521-526, the same float expressions on the same
 * points): side1 = v1 - v0, side2 = v2 - v0, VectorCrossProduct
 * (includes/FloatVector.h:293-296), VectorNormalize. */
void orc_face_normal(const float tv[9], float n[3])
{
    float a[3], b[3];
    for (int q = 0; q < 3; q++) {
        a[q] = tv[3 + q] - tv[q];
        b[q] = tv[6 + q] - tv[q];
    }
    n[0] = a[1] * b[2] - b[1] * a[2];
    n[1] = a[2] * b[0] - b[2] * a[0];
    n[2] = a[0] * b[1] - b[0] * a[1];
    orc_normalize(n);
}

void orc_face_normal_batch(int64_t n, const float *tv, float *out)
{
    for (int64_t i = 0; i < n; i++) orc_face_normal(tv + 9 * i, out + 3 * i);
}

void orc_normalize_batch(int64_t n, const float *v, float *out)
{
    for (int64_t i = 0; i < n; i++) {
        out[3 * i] = v[3 * i];
        out[3 * i + 1] = v[3 * i + 1];
        out[3 * i + 2] = v[3 * i + 2];
        orc_normalize(out + 3 * i);
    }
}

/* Level1SAT list average, src/Object.cpp:3219-3253: face normals summed in bin
 * order (ascending triangle id: ClassifyTessellation walks the triangles in
 * order and AddPointToVoxel appends, :2256-2342, 2705-2748), divided by the
 * (float) count, normalized. */
static void average_normal(const float *fn, int64_t n, float out[3])
{
    float s[3] = {0, 0, 0};
    for (int64_t e = 0; e < n; e++) {
        s[0] += fn[3 * e];
        s[1] += fn[3 * e + 1];
        s[2] += fn[3 * e + 2];
    }
    const float cnt = (float)n;
    out[0] = s[0] / cnt;
    out[1] = s[1] / cnt;
    out[2] = s[2] / cnt;
    orc_normalize(out);
}

void orc_average_normals(int64_t nl, const int64_t *off, const float *fn, float *out)
{
    for (int64_t l = 0; l < nl; l++) average_normal(fn + 3 * off[l], off[l + 1] - off[l], out + 3 * l);
}

static void l1_normal(const float *tris, const int32_t *list, int64_t n, float out[3])
{
    float *fn = (float *)xcalloc((size_t)(n > 0 ? n : 1) * 3, sizeof(float));
    for (int64_t e = 0; e < n; e++) orc_face_normal(tris + 9 * (int64_t)list[e], fn + 3 * e);
    average_normal(fn, n, out);
    free(fn);
}

/* Level1 normals per boundary voxel (l1n[nb*3]) and Level2 normals per Level2
 * voxel (l2n[nb*n2^3*3]; NULL to skip).  Level2: every bin triangle whose SAT
 * hits the Level2 voxel adds its face normal and 1 to a count
 * (AddPointToVoxel, src/Object.cpp:2705-2720) in bin order (the loop nest of
 * :2372-2400); the sum is then divided by the count and normalized
 * (:2404-2421).  Voxels with no hit keep (0, 0, 0). */
int orc_normals(const float *tris, const orc_grid *g, const orc_result *r, int threads, float *l1n,
                float *l2n)
{
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    const int64_t nx = g->n[0], ny = g->n[1];
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t b = 0; b < r->nb; b++) {
        const int64_t v = r->bidx[b];
        l1_normal(tris, r->list_tri + r->list_off[v], r->list_off[v + 1] - r->list_off[v], l1n + 3 * b);
    }
    if (!l2n || g->n2 <= 0) return 0;
    const int n2 = g->n2;
    const int64_t n23 = (int64_t)n2 * n2 * n2;
    float *sx = (float *)xcalloc((size_t)g->n[0] * n2, sizeof(float));
    float *sy = (float *)xcalloc((size_t)g->n[1] * n2, sizeof(float));
    float *sz = (float *)xcalloc((size_t)g->n[2] * n2, sizeof(float));
    orc_l2_centers(g, 0, sx);
    orc_l2_centers(g, 1, sy);
    orc_l2_centers(g, 2, sz);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t b = 0; b < r->nb; b++) {
        const int64_t l1 = r->bidx[b];
        const int64_t k1 = l1 / (nx * ny), ij = l1 - k1 * nx * ny, j1 = ij / nx, i1 = ij % nx;
        float *acc = (float *)xcalloc((size_t)n23 * 4, sizeof(float));
        for (int64_t e = r->list_off[l1]; e < r->list_off[l1 + 1]; e++) {
            const float *tv = tris + 9 * (int64_t)r->list_tri[e];
            float fn[3];
            orc_face_normal(tv, fn);
            for (int p = 0; p < n2; p++)
                for (int q = 0; q < n2; q++)
                    for (int s = 0; s < n2; s++) {
                        const float ctr[3] = {sx[i1 * n2 + p], sy[j1 * n2 + q], sz[k1 * n2 + s]};
                        if (orc_tri_box_overlap(ctr, g->ext2, tv)) {
                            float *a = acc + 4 * (((int64_t)s * n2 + q) * n2 + p);
                            a[0] += fn[0];
                            a[1] += fn[1];
                            a[2] += fn[2];
                            a[3] += 1;
                        }
                    }
        }
        for (int64_t i = 0; i < n23; i++) {
            float o[3] = {0, 0, 0};
            const float cnt = acc[4 * i + 3];
            if (cnt > 0) {
                o[0] = acc[4 * i + 0] / cnt;
                o[1] = acc[4 * i + 1] / cnt;
                o[2] = acc[4 * i + 2] / cnt;
                orc_normalize(o);
            }
            float *d = l2n + 3 * (b * n23 + i);
            d[0] = o[0];
            d[1] = o[1];
            d[2] = o[2];
        }
        free(acc);
    }
    free(sx);
    free(sy);
    free(sz);
    return 0;
}

/* .raw normal byte, src/Object.cpp:2940-2942 and 3032-3034, 3047-3049:
 * outputDType(n * normScale + normBias) with normScale = 256.0/3.0 and
 * normBias = 127.0 held in floats. */
uint8_t orc_quantize_normal(float v)
{
    const float scale = (float)(256.0 / 3.0), bias = 127.0f;
    const float y = v * scale;
    return (uint8_t)(y + bias);
}

void orc_quantize_normals(int64_t n, const float *v, uint8_t *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = orc_quantize_normal(v[i]);
}

void orc_result_free(orc_result *r)
{
    free(r->l1);
    free(r->prefix);
    free(r->bidx);
    free(r->l2);
    free(r->list_off);
    free(r->list_tri);
    memset(r, 0, sizeof(*r));
}

/* src/Object.cpp:2934-3075 (in/out, prefix and config files; states are
 * written as uint8(state * 127.0f)). */
int orc_save_raw(const orc_grid *g, const orc_result *r, int obj_id, const char *dir)
{
    char path[4096];
    FILE *f;
    snprintf(path, sizeof path, "%s/Obj%dVoxelConfig.txt", dir, obj_id);
    if (!(f = fopen(path, "w"))) return 1;
    fprintf(f, "Obj%d\n", obj_id);
    fprintf(f, "%g\t%g\t%g\n", g->bmin[0], g->bmin[1], g->bmin[2]);
    fprintf(f, "%g\t%g\t%g\n", g->bmax[0], g->bmax[1], g->bmax[2]);
    fprintf(f, "%d\t%d\t%d\n", g->n[0], g->n[1], g->n[2]);
    fprintf(f, "%g\t%g\t%g\n", g->g[0], g->g[1], g->g[2]);
    fprintf(f, "%lld\n%lld\n", (long long)r->l1_inside, (long long)r->l1_boundary);
    if (g->n2 > 0) {
        fprintf(f, "%d\t%d\t%d\n", g->n2, g->n2, g->n2);
        fprintf(f, "%g\t%g\t%g\n", g->g2[0], g->g2[1], g->g2[2]);
        fprintf(f, "%lld\n%lld\n", (long long)r->l2_inside, (long long)r->l2_boundary);
    }
    fclose(f);
    uint8_t *buf = (uint8_t *)xcalloc((size_t)r->n1, 1);
    for (int64_t v = 0; v < r->n1; v++) buf[v] = (uint8_t)((float)r->l1[v] * 127.0f);
    snprintf(path, sizeof path, "%s/Obj%dLevel1InOut.raw", dir, obj_id);
    if (!(f = fopen(path, "wb"))) { free(buf); return 1; }
    fwrite(buf, 1, (size_t)r->n1, f);
    fclose(f);
    free(buf);
    if (g->n2 > 0) {
        snprintf(path, sizeof path, "%s/Obj%dLevel1BoundaryPrefixSum.raw", dir, obj_id);
        if (!(f = fopen(path, "wb"))) return 1;
        fwrite(r->prefix, sizeof(int32_t), (size_t)r->n1, f);
        fclose(f);
        int64_t n = r->nb * (int64_t)g->n2 * g->n2 * g->n2;
        buf = (uint8_t *)xcalloc((size_t)n, 1);
        for (int64_t v = 0; v < n; v++) buf[v] = (uint8_t)((float)r->l2[v] * 127.0f);
        snprintf(path, sizeof path, "%s/Obj%dLevel2InOut.raw", dir, obj_id);
        if (!(f = fopen(path, "wb"))) { free(buf); return 1; }
        fwrite(buf, 1, (size_t)n, f);
        fclose(f);
        free(buf);
    }
    return 0;
}

/* The two normal files of SaveVoxelization (src/Object.cpp:3024-3060):
 * Obj<id>Level1Normal.raw, uint8[n1][3], the boundary voxels' normals (l1n,
 * bidx order) and 0 elsewhere, and Obj<id>Level2Normal.raw,
 * uint8[nb*n2^3][3], each component uint8(n*256/3 + 127). */
int orc_save_raw_normals(const orc_grid *g, const orc_result *r, const float *l1n, const float *l2n, int obj_id,
                         const char *dir)
{
    char path[4096];
    FILE *f;
    uint8_t *buf = (uint8_t *)xcalloc((size_t)r->n1 * 3, 1);
    memset(buf, orc_quantize_normal(0.0f), (size_t)r->n1 * 3);
    for (int64_t b = 0; b < r->nb; b++)
        for (int q = 0; q < 3; q++) buf[(int64_t)r->bidx[b] * 3 + q] = orc_quantize_normal(l1n[b * 3 + q]);
    snprintf(path, sizeof path, "%s/Obj%dLevel1Normal.raw", dir, obj_id);
    if (!(f = fopen(path, "wb"))) { free(buf); return 1; }
    fwrite(buf, 1, (size_t)r->n1 * 3, f);
    fclose(f);
    free(buf);
    if (g->n2 > 0) {
        int64_t n = r->nb * (int64_t)g->n2 * g->n2 * g->n2 * 3;
        buf = (uint8_t *)xcalloc((size_t)n, 1);
        for (int64_t v = 0; v < n; v++) buf[v] = orc_quantize_normal(l2n[v]);
        snprintf(path, sizeof path, "%s/Obj%dLevel2Normal.raw", dir, obj_id);
        if (!(f = fopen(path, "wb"))) { free(buf); return 1; }
        fwrite(buf, 1, (size_t)n, f);
        fclose(f);
        free(buf);
    }
    return 0;
}
