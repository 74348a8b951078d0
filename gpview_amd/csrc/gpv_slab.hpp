// gpv_slab.hpp -- conservative plane-slab filter for the SAT's plane predicate.
//
// The SAT (src/TriBoxIntersection.cpp:135-210) is a conjunction; its plane
// predicate (planeBoxOverlap, :64-86) is evaluated on float values relative
// to the box center: v_i = fl(t_i - c), e0 = fl(v1 - v0), e1 = fl(v2 - v1),
// n = fl(e0 x e1), d1 = fl(n . vmin), d2 = fl(n . vmax) with vmin/vmax chosen
// by the sign of n.  In exact arithmetic, with N = (t1 - t0) x (t2 - t1) and
// s = N . (c - t0), R = sum h_q |N_q|: d1 = s - R and d2 = s + R.  Rounding
// moves d1, d2 by at most M (bounded below per term), so the float predicate
// rejects for certain when |s| > R + M.  Boxes outside that slab can be
// skipped without evaluating the SAT; every other box still runs the full
// float SAT, so results are unchanged.
//
// M, per axis q, with A_iq bounding |t_iq - c_q| over the candidate centers:
//   D0_q = U (A0 + A1 + |E0|)      |e0_q - E0_q|      (E0 = t1 - t0)
//   D1_q = U (A1 + A2 + |E1|)      |e1_q - E1_q|      (E1 = t2 - t1)
//   Nerr = cross(D, F) terms + 2U (F0 F1 terms)        |n_q - N_q|
//   M = 3U sum |n| V + sum Nerr V + sum |N| Verr + sum 2 h Nerr
// with V_q = h_q + A0_q and Verr_q = U (2 A0_q + h_q); every term carries a
// 1% cushion and the whole M a factor kSlabSafety.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace gpv {

constexpr double kSlabSafety = 4.0;

struct PlaneSlab {
    double nx, ny, nz;   // exact normal of the float triangle
    double t0x, t0y, t0z;
    double lim;          // keep a box iff |N . (c - t0)| <= lim
};

// clo/chi: smallest and largest candidate center per axis; h: half extents.
__host__ __device__ inline PlaneSlab plane_slab(const float* tv, float hx, float hy, float hz,
                                                const double clo[3], const double chi[3])
{
    const double U = 5.9604644775390625e-08 * 1.01;
    const double h[3] = {hx, hy, hz};
    double A[3][3], E0[3], E1[3];
    for (int q = 0; q < 3; ++q) {
        for (int i = 0; i < 3; ++i) {
            const double t = tv[3 * i + q];
            A[i][q] = fmax(fabs(t - clo[q]), fabs(t - chi[q]));
        }
        E0[q] = (double)tv[3 + q] - (double)tv[q];
        E1[q] = (double)tv[6 + q] - (double)tv[3 + q];
    }
    double D0[3], D1[3], F0[3], F1[3];
    for (int q = 0; q < 3; ++q) {
        D0[q] = U * (A[0][q] + A[1][q] + fabs(E0[q]));
        D1[q] = U * (A[1][q] + A[2][q] + fabs(E1[q]));
        F0[q] = fabs(E0[q]) + D0[q];
        F1[q] = fabs(E1[q]) + D1[q];
    }
    PlaneSlab s;
    s.nx = E0[1] * E1[2] - E0[2] * E1[1];
    s.ny = E0[2] * E1[0] - E0[0] * E1[2];
    s.nz = E0[0] * E1[1] - E0[1] * E1[0];
    s.t0x = tv[0];
    s.t0y = tv[1];
    s.t0z = tv[2];
    const double N[3] = {s.nx, s.ny, s.nz};
    // component q of the cross product uses axes a = q+1, b = q+2
    double R = 0.0, M = 0.0;
    for (int q = 0; q < 3; ++q) {
        const int a = (q + 1) % 3, b = (q + 2) % 3;
        const double ce = D0[a] * F1[b] + fabs(E0[a]) * D1[b] + D0[b] * F1[a] + fabs(E0[b]) * D1[a];
        const double rc = 2.0 * U * (F0[a] * F1[b] + F0[b] * F1[a]);
        const double nerr = (ce + rc) * 1.01 + 1e-12 * (fabs(E0[a] * E1[b]) + fabs(E0[b] * E1[a]));
        const double nb = fabs(N[q]) + nerr;
        const double V = h[q] + A[0][q] * 1.01;
        const double verr = U * (2.0 * A[0][q] + h[q]);
        R += h[q] * fabs(N[q]);
        M += 3.0 * U * nb * V + nerr * V + fabs(N[q]) * verr + 2.0 * h[q] * nerr;
    }
    // evaluation slop of s itself (double products and sums of these sizes)
    double cs = 0.0;
    for (int q = 0; q < 3; ++q) cs += fabs(N[q]) * (fmax(fabs(clo[q]), fabs(chi[q])) + fabs((double)tv[q]));
    s.lim = R + kSlabSafety * M + 1e-12 * cs;
    return s;
}

// Float form of a slab for centers within [clo, chi]: the normal rounded to
// float and the bound grown by the float evaluation error of
// fl(fl(fl(nx*(cx - t0x)) + fl(ny*(cy - t0y))) + fl(nz*(cz - t0z))), at most
// 6U sum |N_q| max|c_q - t0_q| (1U for the rounded normal, 1U per subtraction,
// product and sum), taken as 8U and rounded up to a float.
struct FloatSlab {
    float nx, ny, nz, lim;
};

__host__ __device__ inline FloatSlab float_slab(const PlaneSlab& sl, const float* tv, const double clo[3],
                                                const double chi[3])
{
    double sa = 0.0;
    sa += fabs(sl.nx) * fmax(fabs(clo[0] - tv[0]), fabs(chi[0] - tv[0]));
    sa += fabs(sl.ny) * fmax(fabs(clo[1] - tv[1]), fabs(chi[1] - tv[1]));
    sa += fabs(sl.nz) * fmax(fabs(clo[2] - tv[2]), fabs(chi[2] - tv[2]));
    const double limd = sl.lim + 8.0 * 5.9604644775390625e-08 * sa;
    float limf = (float)limd;
    if ((double)limf < limd) limf = nextafterf(limf, INFINITY);
    return FloatSlab{(float)sl.nx, (float)sl.ny, (float)sl.nz, limf};
}

// the float test every caller of a FloatSlab evaluates, in this order
__host__ __device__ inline bool float_slab_keeps(const FloatSlab& f, const float* tv, float cx, float cy, float cz)
{
    const float a = f.nx * (cx - tv[0]) + f.ny * (cy - tv[1]);
    const float d = a + f.nz * (cz - tv[2]);
    return fabsf(d) <= f.lim;
}

__host__ __device__ inline double slab_dist(const PlaneSlab& s, double cx, double cy, double cz)
{
    return s.nx * (cx - s.t0x) + s.ny * (cy - s.t0y) + s.nz * (cz - s.t0z);
}

// Layers [za, zb) of the column with center (cx, cy) that a slab (normal n,
// bound lim, anchored at vertex tv[0..2]) can keep: |n.(c - t0)| <= lim is
// linear in the center's z, so the kept centers form an interval, located on
// the monotone center table c1z[z0, z1) itself and widened by one layer on
// each side.  Every kept layer still runs the full SAT; layers outside are
// rejected by it.
__host__ __device__ inline void slab_layers(double nx, double ny, double nz, double lim, float cx, float cy,
                                            const float* tv, const float* c1z, int z0, int z1, int& za, int& zb)
{
    za = z0;
    zb = z1;
    const double A = nx * ((double)cx - (double)tv[0]) + ny * ((double)cy - (double)tv[1]);
    if (nz == 0.0) {
        if (fabs(A) > lim) zb = za;
        return;
    }
    double lo = (-lim - A) / nz, hi = (lim - A) / nz;
    if (lo > hi) {
        const double t = lo;
        lo = hi;
        hi = t;
    }
    lo += (double)tv[2];
    hi += (double)tv[2];
    if (!(lo <= hi)) return;   // NaN: keep every layer
    int a = z0, b = z1;        // first layer with center >= lo
    while (b > a) {
        const int m = (a + b) >> 1;
        if ((double)c1z[m] < lo) a = m + 1; else b = m;
    }
    int c = a, d = z1;         // first layer with center > hi
    while (d > c) {
        const int m = (c + d) >> 1;
        if ((double)c1z[m] <= hi) c = m + 1; else d = m;
    }
    za = a - 1 > z0 ? a - 1 : z0;
    zb = c + 1 < z1 ? c + 1 : z1;
}

// ---------------------------------------------------------------------------
// Level2 parity: the +z ray test's t > 1e-6 decided for a whole voxel.
// For a column record (edges e1, e2, 1/det inv, vertex v1) the parity kernels
// evaluate per ray origin o, in float (src/TriRayIntersection.cpp:112-122):
//   tx = o.x - v1x, ty = o.y - v1y, tz = o.z - v1z
//   q  = (ty e1z - tz e1y, tz e1x - tx e1z, tx e1y - ty e1x)
//   s  = (e2x qx + e2y qy) + e2z qz,  t = s * inv,  pass iff t >= eps.
// Rounding is monotone, so for every origin of a voxel T_a lies in
// [tl_a, th_a] = [fl(olo_a - v1_a), fl(ohi_a - v1_a)].  Let n = e1 x e2,
// N_a = |e1_b e2_c| + |e1_c e2_b| (b, c the other axes), M_a = max(|tl_a|,
// |th_a|) and B = sum_a M_a N_a.  S = e2 . (T x e1) = T . n is the exact value
// on the float T, affine in T, so g S (g = sign(inv)) ranges over
// [sum_a min(g n_a tl_a, g n_a th_a), sum_a max(...)] on the box.
//   - each q_a is within 2U of its two products, and the dot product adds 3U
//     per term: |s - S| <= (5U + 6U^2) B (the bound K of the products is B);
//   - the float n_a is within 2U N_a, each float product adds U, the sum of
//     three terms 2U: the computed range ends are within 6.0001U B.
// So g s lies in [lo - E, hi + E] with E = 16U B_f (B_f the float B, itself
// within 5U of B) plus 2^-100 for gradual underflow.  t = fl(g s |inv|):
// pass for every origin when fl(fl(lo - E) |inv|) >= eps (1 + 2^-18), fail
// for every origin when fl(hi + E) <= 0 or fl(fl(hi + E) |inv|) <= eps
// (1 - 2^-18); the 2^-18 margins absorb the last few roundings.  Returns +1,
// -1, or 0 when undecided.  `slack` scales E (1 in the kernels; the tests
// lower it to show the terms are needed).  All in float: few registers.
// ---------------------------------------------------------------------------
__host__ __device__ inline void t_axis(float n, float tl, float th, float nabs, float& lo, float& hi, float& B)
{
    const float p = n * tl, q = n * th;
    lo += fminf(p, q);
    hi += fmaxf(p, q);
    B += fmaxf(fabsf(tl), fabsf(th)) * nabs;
}

__host__ __device__ inline int t_decided(float e1x, float e1y, float e1z, float e2x, float e2y, float e2z, float inv,
                                         float v1x, float v1y, float v1z, float oxlo, float oxhi, float oylo,
                                         float oyhi, float ozlo, float ozhi, float eps, float slack = 1.0f)
{
    const float g = inv > 0.0f ? 1.0f : -1.0f;
    float lo = 0.0f, hi = 0.0f, B = 0.0f;
    t_axis(g * (e1y * e2z - e1z * e2y), oxlo - v1x, oxhi - v1x, fabsf(e1y * e2z) + fabsf(e1z * e2y), lo, hi, B);
    t_axis(g * (e1z * e2x - e1x * e2z), oylo - v1y, oyhi - v1y, fabsf(e1z * e2x) + fabsf(e1x * e2z), lo, hi, B);
    t_axis(g * (e1x * e2y - e1y * e2x), ozlo - v1z, ozhi - v1z, fabsf(e1x * e2y) + fabsf(e1y * e2x), lo, hi, B);
    if (!(B <= 3.0e38f) || !(fabsf(lo) <= 3.0e38f) || !(fabsf(hi) <= 3.0e38f)) return 0;
    const float E = slack * (B * 9.5367431640625e-07f + 7.888609052210118e-31f);   // 16U B + 2^-100
    const float ai = fabsf(inv);
    const float l = lo - E, m = hi + E;
    if (l > 0.0f && l * ai >= eps * 1.0000038f) return 1;
    if (m <= 0.0f || m * ai <= eps * 0.9999962f) return -1;
    return 0;
}

}  // namespace gpv
