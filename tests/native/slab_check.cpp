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
