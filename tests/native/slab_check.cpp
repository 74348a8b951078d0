// Host-side harness for the plane-slab filter (gpview_amd/csrc/gpv_slab.hpp):
// tests/test_slab_filter.py checks that no box the reference SAT accepts is
// ever outside the slab.  Test tooling only.
#include <stdint.h>
#include "gpv_slab.hpp"

extern "C" void slab_keep_batch(int64_t n, const float* tv, const float* c, const float* h, uint8_t* keep)
{
    for (int64_t i = 0; i < n; ++i) {
        const double lo[3] = {c[3 * i], c[3 * i + 1], c[3 * i + 2]};
        const gpv::PlaneSlab s = gpv::plane_slab(tv + 9 * i, h[3 * i], h[3 * i + 1], h[3 * i + 2], lo, lo);
        const double d = gpv::slab_dist(s, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
        keep[i] = (uint8_t)(fabs(d) <= s.lim);
    }
}

// Level1 layer intervals: triangle i with index box lo[3i..], hi[3i..]
// (inclusive) writes, for every column of its box in (q, p) order from
// off[i], the kept layer interval [za, zb) of gpv::slab_layers.
extern "C" void slab_layers_batch(int64_t n, const float* tv, const int32_t* lo, const int32_t* hi,
                                  const float* c1x, const float* c1y, const float* c1z, const float* h,
                                  const int64_t* off, int32_t* za, int32_t* zb)
{
    for (int64_t i = 0; i < n; ++i) {
        const int32_t* l = lo + 3 * i;
        const int32_t* u = hi + 3 * i;
        const double clo[3] = {c1x[l[0]], c1y[l[1]], c1z[l[2]]};
        const double chi[3] = {c1x[u[0]], c1y[u[1]], c1z[u[2]]};
        const gpv::PlaneSlab s = gpv::plane_slab(tv + 9 * i, h[0], h[1], h[2], clo, chi);
        int64_t k = off[i];
        for (int q = l[1]; q <= u[1]; ++q)
            for (int p = l[0]; p <= u[0]; ++p, ++k) {
                int a, b;
                gpv::slab_layers(s.nx, s.ny, s.nz, s.lim, c1x[p], c1y[q], tv + 9 * i, c1z, l[2], u[2] + 1, a, b);
                za[k] = a;
                zb[k] = b;
            }
    }
}

// The Level2 float slab: slab and bound over the centers range [clo, chi] of
// each case (per case 6 doubles), the float test at center c.
extern "C" void float_slab_keep_batch(int64_t n, const float* tv, const float* c, const float* h,
                                      const double* lohi, uint8_t* keep)
{
    for (int64_t i = 0; i < n; ++i) {
        const double* r = lohi + 6 * i;
        const double lo[3] = {r[0], r[1], r[2]}, hi[3] = {r[3], r[4], r[5]};
        const gpv::PlaneSlab s = gpv::plane_slab(tv + 9 * i, h[3 * i], h[3 * i + 1], h[3 * i + 2], lo, hi);
        const gpv::FloatSlab f = gpv::float_slab(s, tv + 9 * i, lo, hi);
        keep[i] = (uint8_t)gpv::float_slab_keeps(f, tv + 9 * i, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    }
}

// Level2 parity: gpv::t_decided for the column record of triangle i (edges,
// det and 1/det in the float expressions of the +z Moller-Trumbore test,
// src/TriRayIntersection.cpp:86-101) over the origin box
// box[6i..6i+6) = (oxlo, oxhi, oylo, oyhi, ozlo, ozhi).  dec = 2 where the
// record's det fails (no ray can hit).
extern "C" void t_decided_batch(int64_t n, const float* tv, const float* box, float eps, double slack, int8_t* dec)
{
    for (int64_t i = 0; i < n; ++i) {
        const float* t = tv + 9 * i;
        const float e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
        const float e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
        const float px = 0.0f * e2z - e2y, py = e2x - 0.0f * e2z, pz = 0.0f * e2y - 0.0f * e2x;
        const float det = e1x * px + e1y * py + e1z * pz;
        if ((double)det > -1e-6 && (double)det < 1e-6) {
            dec[i] = 2;
            continue;
        }
        const float inv = 1.f / det;
        const float* b = box + 6 * i;
        dec[i] = (int8_t)gpv::t_decided(e1x, e1y, e1z, e2x, e2y, e2z, inv, t[0], t[1], t[2], b[0], b[1], b[2], b[3],
                                        b[4], b[5], eps, slack);
    }
}
