// gpv_device.hpp -- device-side geometry predicates for the gfx950 voxelizer.
//
// Exactness contract: every expression below is evaluated in IEEE binary32
// with the operand order of the reference's CPU predicates
// (src/TriBoxIntersection.cpp:64-210, src/TriRayIntersection.cpp:78-131).
// The library is compiled with -ffp-contract=off and without fast-math, so no
// multiply/add pair is fused, nothing is reassociated, f32 denormals are kept
// and 1.f/det is the correctly rounded IEEE quotient.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpv {

// (double)x > 1e-6  <=>  x >= kEpsAbove   (smallest float above 1e-6)
// (double)x < 1e-6  <=>  x <= kEpsBelow   (largest float below 1e-6)
// Both are checked against the double comparison on the host by the tests.
constexpr uint32_t kEpsAboveBits = 0x358637BEu;
constexpr uint32_t kEpsBelowBits = 0x358637BDu;

__device__ __forceinline__ float eps_above() { return __uint_as_float(kEpsAboveBits); }
__device__ __forceinline__ float eps_below() { return __uint_as_float(kEpsBelowBits); }

// ---------------------------------------------------------------------------
// SAT triangle/box overlap.  The reference returns the conjunction of 13
// predicates (9 edge axes, 3 box axes, the triangle plane); the order of
// evaluation cannot change a conjunction, so the box axes run first because
// they reject most often in a sweep.  Each predicate is the reference's float
// expression unchanged, its two projections passed in the order the
// reference's macro compares them: (p0, p2) for X01/Y02, (p0, p1) for
// X2/Y1/Z0, (p2, p1) for Z12.  The order decides which of a NaN pair
// becomes the minimum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool sat_separated(float pa, float pb, float rad)
{
    float mn, mx;
    if (pa < pb) { mn = pa; mx = pb; } else { mn = pb; mx = pa; }
    return (mn > rad) | (mx < -rad);
}

// The three edge-axis predicates along e x z (the third AXISTEST of each
// edge) read only the x and y components of v_i = tri_i - center and of the
// edges, which depend on (cx, cy) alone: they hold or fail for a whole column
// of boxes.  Same float expressions as in tri_box_overlap.
__device__ __forceinline__ bool xy_edges_ok(float cx, float cy, float hx, float hy, const float* __restrict__ tv)
{
    const float v0x = tv[0] - cx, v0y = tv[1] - cy;
    const float v1x = tv[3] - cx, v1y = tv[4] - cy;
    const float v2x = tv[6] - cx, v2y = tv[7] - cy;
    const float e0x = v1x - v0x, e0y = v1y - v0y;
    const float e1x = v2x - v1x, e1y = v2y - v1y;
    const float e2x = v0x - v2x, e2y = v0y - v2y;
    if (sat_separated(e0y * v2x - e0x * v2y, e0y * v1x - e0x * v1y, fabsf(e0y) * hx + fabsf(e0x) * hy)) return false;
    if (sat_separated(e1y * v0x - e1x * v0y, e1y * v1x - e1x * v1y, fabsf(e1y) * hx + fabsf(e1x) * hy)) return false;
    if (sat_separated(e2y * v2x - e2x * v2y, e2y * v1x - e2x * v1y, fabsf(e2y) * hx + fabsf(e2x) * hy)) return false;
    return true;
}

// kAxes = false skips the three box-axis predicates; only for callers that
// have already established them with sat_axis_ok on the same center.
// kZEdges = false skips the three e x z edge predicates; only for callers that
// have established them with xy_edges_ok on the same (cx, cy).
template <bool kAxes = true, bool kZEdges = true>
__device__ __forceinline__ bool tri_box_overlap(float cx, float cy, float cz,
                                                float hx, float hy, float hz,
                                                const float* __restrict__ tv)
{
    const float v0x = tv[0] - cx, v0y = tv[1] - cy, v0z = tv[2] - cz;
    const float v1x = tv[3] - cx, v1y = tv[4] - cy, v1z = tv[5] - cz;
    const float v2x = tv[6] - cx, v2y = tv[7] - cy, v2z = tv[8] - cz;

    // box axes (FINDMINMAX, src/TriBoxIntersection.cpp:191-200)
    if (kAxes) {
        float mn = v0x, mx = v0x;
        if (v1x < mn) mn = v1x;
        if (v1x > mx) mx = v1x;
        if (v2x < mn) mn = v2x;
        if (v2x > mx) mx = v2x;
        if ((mn > hx) | (mx < -hx)) return false;
    }
    if (kAxes) {
        float mn = v0y, mx = v0y;
        if (v1y < mn) mn = v1y;
        if (v1y > mx) mx = v1y;
        if (v2y < mn) mn = v2y;
        if (v2y > mx) mx = v2y;
        if ((mn > hy) | (mx < -hy)) return false;
    }
    if (kAxes) {
        float mn = v0z, mx = v0z;
        if (v1z < mn) mn = v1z;
        if (v1z > mx) mx = v1z;
        if (v2z < mn) mn = v2z;
        if (v2z > mx) mx = v2z;
        if ((mn > hz) | (mx < -hz)) return false;
    }

    const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
    const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
    const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;

    // triangle plane, n = e0 x e1 (src/TriBoxIntersection.cpp:64-86, 205-207)
    {
        const float nx = e0y * e1z - e0z * e1y;
        const float ny = e0z * e1x - e0x * e1z;
        const float nz = e0x * e1y - e0y * e1x;
        float vminx, vmaxx, vminy, vmaxy, vminz, vmaxz;
        if (nx > 0.0f) { vminx = -hx - v0x; vmaxx = hx - v0x; } else { vminx = hx - v0x; vmaxx = -hx - v0x; }
        if (ny > 0.0f) { vminy = -hy - v0y; vmaxy = hy - v0y; } else { vminy = hy - v0y; vmaxy = -hy - v0y; }
        if (nz > 0.0f) { vminz = -hz - v0z; vmaxz = hz - v0z; } else { vminz = hz - v0z; vmaxz = -hz - v0z; }
        if (nx * vminx + ny * vminy + nz * vminz > 0.0f) return false;
        if (!(nx * vmaxx + ny * vmaxy + nz * vmaxz >= 0.0f)) return false;
    }

    // nine edge axes (AXISTEST_*, src/TriBoxIntersection.cpp:90-133, 163-182)
    float fx = fabsf(e0x), fy = fabsf(e0y), fz = fabsf(e0z);
    if (sat_separated(e0z * v0y - e0y * v0z, e0z * v2y - e0y * v2z, fz * hy + fy * hz)) return false;
    if (sat_separated(-e0z * v0x + e0x * v0z, -e0z * v2x + e0x * v2z, fz * hx + fx * hz)) return false;
    if (kZEdges && sat_separated(e0y * v2x - e0x * v2y, e0y * v1x - e0x * v1y, fy * hx + fx * hy)) return false;
    fx = fabsf(e1x); fy = fabsf(e1y); fz = fabsf(e1z);
    if (sat_separated(e1z * v0y - e1y * v0z, e1z * v2y - e1y * v2z, fz * hy + fy * hz)) return false;
    if (sat_separated(-e1z * v0x + e1x * v0z, -e1z * v2x + e1x * v2z, fz * hx + fx * hz)) return false;
    if (kZEdges && sat_separated(e1y * v0x - e1x * v0y, e1y * v1x - e1x * v1y, fy * hx + fx * hy)) return false;
    fx = fabsf(e2x); fy = fabsf(e2y); fz = fabsf(e2z);
    if (sat_separated(e2z * v0y - e2y * v0z, e2z * v1y - e2y * v1z, fz * hy + fy * hz)) return false;
    if (sat_separated(-e2z * v0x + e2x * v0z, -e2z * v1x + e2x * v1z, fz * hx + fx * hz)) return false;
    if (kZEdges && sat_separated(e2y * v2x - e2x * v2y, e2y * v1x - e2x * v1y, fy * hx + fx * hy)) return false;
    return true;
}

// Box-axis part of the SAT on one axis: true when the box at center c (half
// size h) is NOT separated from the triangle along that axis.  Exactly the
// float test tri_box_overlap runs, so a range pruned with it loses nothing.
__device__ __forceinline__ bool sat_axis_ok(float a0, float a1, float a2, float c, float h)
{
    const float d0 = a0 - c, d1 = a1 - c, d2 = a2 - c;
    float mn = d0, mx = d0;
    if (d1 < mn) mn = d1;
    if (d1 > mx) mx = d1;
    if (d2 < mn) mn = d2;
    if (d2 > mx) mx = d2;
    return !((mn > h) | (mx < -h));
}

// ---------------------------------------------------------------------------
// Moller-Trumbore for the +z ray D = (0,0,1), split at the column seam: with
// that D every O.z term of P, det, u and v is multiplied by an exact zero, so
// the det/u/v decision belongs to the (O.x, O.y) column and only t to the
// voxel.  The float expressions are those of src/TriRayIntersection.cpp:86-122
// with D's components substituted; the exact-zero products only add signed
// zeros, which change no comparison.
// ---------------------------------------------------------------------------
struct MtTri {
    float e1x, e1y, e1z, e2x, e2y, e2z;
    float det, inv_det;
};

// det of the +z ray; false if the ray test rejects it for every origin
__device__ __forceinline__ bool mt_setup(const float* __restrict__ tv, MtTri& m)
{
    m.e1x = tv[3] - tv[0]; m.e1y = tv[4] - tv[1]; m.e1z = tv[5] - tv[2];
    m.e2x = tv[6] - tv[0]; m.e2y = tv[7] - tv[1]; m.e2z = tv[8] - tv[2];
    // P = D x e2 = (0*e2z - 1*e2y, 1*e2x - 0*e2z, 0*e2y - 0*e2x)
    const float px = 0.0f * m.e2z - m.e2y;
    const float py = m.e2x - 0.0f * m.e2z;
    const float pz = 0.0f * m.e2y - 0.0f * m.e2x;
    m.det = m.e1x * px + m.e1y * py + m.e1z * pz;
    const float eb = eps_below();
    if ((m.det >= -eb) & (m.det <= eb)) return false;   // (double)det in (-1e-6, 1e-6)
    m.inv_det = 1.f / m.det;
    return true;
}

// u/v acceptance for the column through (ox, oy); returns T.x, T.y for t.
__device__ __forceinline__ bool mt_column(const MtTri& m, float v1x, float v1y, float ox, float oy,
                                          float& tx, float& ty)
{
    tx = ox - v1x;
    ty = oy - v1y;
    const float px = 0.0f * m.e2z - m.e2y;
    const float py = m.e2x - 0.0f * m.e2z;
    const float u = (tx * px + ty * py) * m.inv_det;
    if ((u < 0.f) | (u > 1.f)) return false;
    const float qz = tx * m.e1y - ty * m.e1x;
    const float v = qz * m.inv_det;
    if ((v < 0.f) | (u + v > 1.f)) return false;
    return true;
}

// t > 1e-6 for the voxel at height oz (v1z = V1.z)
__device__ __forceinline__ bool mt_t_hit(float e1x, float e1y, float e1z, float e2x, float e2y,
                                         float e2z, float inv_det, float tx, float ty, float v1z,
                                         float oz)
{
    const float tz = oz - v1z;
    const float qx = ty * e1z - tz * e1y;
    const float qy = tz * e1x - tx * e1z;
    const float qz = tx * e1y - ty * e1x;
    const float t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
    return t >= eps_above();
}

// ---------------------------------------------------------------------------
// wave helpers (64-wide)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanes_below_mask()
{
    const unsigned l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace gpv
