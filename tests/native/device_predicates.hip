// Test-only harness: runs the device predicates of gpview_amd/csrc/gpv_device.hpp
// on the GPU over batches of golden cases, so tests/test_gpu_predicates.py can
// compare them bit for bit with the reference's own TriBoxOverlap and
// triangle_ray_intersection outputs (tests/golden/{sat,mt}_cases.npz,
// tests/golden/sat_nonfinite.npz).
// Built with the product's flags (no contraction, no fast-math, packed FP32
// off).  Test tooling only: the product never loads it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpv_device.hpp"

using namespace gpv;

// out[4*i + k]: k = 0 the full SAT; 1 the box axes by sat_axis_ok then the
// SAT without them; 2 xy_edges_ok then the SAT without the e x z edges (the
// Level2 queue for sub-boxes not known to pass the box axes); 3 sat_axis_ok
// and xy_edges_ok then the SAT without either (the Level2 queue for
// sub-boxes inside the box-axis masks).
__global__ void k_sat_variants(const float* __restrict__ c, const float* __restrict__ h,
                               const float* __restrict__ tv, int64_t n, uint8_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t[9];
    for (int k = 0; k < 9; ++k) t[k] = tv[9 * i + k];
    const float cx = c[3 * i], cy = c[3 * i + 1], cz = c[3 * i + 2];
    const float hx = h[3 * i], hy = h[3 * i + 1], hz = h[3 * i + 2];
    const bool full = tri_box_overlap<true, true>(cx, cy, cz, hx, hy, hz, t);
    const bool axes = sat_axis_ok(t[0], t[3], t[6], cx, hx) && sat_axis_ok(t[1], t[4], t[7], cy, hy) &&
                      sat_axis_ok(t[2], t[5], t[8], cz, hz);
    const bool zed = xy_edges_ok(cx, cy, hx, hy, t);
    const bool v1 = axes && tri_box_overlap<false, true>(cx, cy, cz, hx, hy, hz, t);
    const bool v2 = zed && tri_box_overlap<true, false>(cx, cy, cz, hx, hy, hz, t);
    const bool v3 = axes && zed && tri_box_overlap<false, false>(cx, cy, cz, hx, hy, hz, t);
    out[4 * i + 0] = full;
    out[4 * i + 1] = v1;
    out[4 * i + 2] = v2;
    out[4 * i + 3] = v3;
}

// +z ray from origin o: mt_setup (det), mt_column (u, v), mt_t_hit (t)
__global__ void k_mt_zray(const float* __restrict__ tv, const float* __restrict__ o, int64_t n,
                          uint8_t* __restrict__ hit)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t[9];
    for (int k = 0; k < 9; ++k) t[k] = tv[9 * i + k];
    MtTri m;
    bool r = mt_setup(t, m);
    float tx = 0.0f, ty = 0.0f;
    if (r) r = mt_column(m, t[0], t[1], o[3 * i], o[3 * i + 1], tx, ty);
    if (r) r = mt_t_hit(m.e1x, m.e1y, m.e1z, m.e2x, m.e2y, m.e2z, m.inv_det, tx, ty, t[2], o[3 * i + 2]);
    hit[i] = r;
}

template <typename T>
static int up(const T* h, size_t n, T** d)
{
    if (hipMalloc((void**)d, n * sizeof(T) + 16) != hipSuccess) return 1;
    return hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess;
}

extern "C" int dp_sat(int64_t n, const float* c, const float* h, const float* tv, uint8_t* out)
{
    float *dc = nullptr, *dh = nullptr, *dt = nullptr;
    uint8_t* dout = nullptr;
    int e = up(c, 3 * n, &dc) | up(h, 3 * n, &dh) | up(tv, 9 * n, &dt);
    if (!e) e = hipMalloc((void**)&dout, 4 * n + 16) != hipSuccess;
    if (!e && n > 0) {
        k_sat_variants<<<(unsigned)((n + 255) / 256), 256>>>(dc, dh, dt, n, dout);
        e = hipGetLastError() != hipSuccess;
    }
    if (!e) e = hipMemcpy(out, dout, 4 * n, hipMemcpyDeviceToHost) != hipSuccess;
    (void)hipFree(dc);
    (void)hipFree(dh);
    (void)hipFree(dt);
    (void)hipFree(dout);
    return e;
}

extern "C" int dp_mt(int64_t n, const float* tv, const float* o, uint8_t* hit)
{
    float *dt = nullptr, *dn = nullptr;
    uint8_t* dh = nullptr;
    int e = up(tv, 9 * n, &dt) | up(o, 3 * n, &dn);
    if (!e) e = hipMalloc((void**)&dh, n + 16) != hipSuccess;
    if (!e && n > 0) {
        k_mt_zray<<<(unsigned)((n + 255) / 256), 256>>>(dt, dn, n, dh);
        e = hipGetLastError() != hipSuccess;
    }
    if (!e) e = hipMemcpy(hit, dh, n, hipMemcpyDeviceToHost) != hipSuccess;
    (void)hipFree(dt);
    (void)hipFree(dn);
    (void)hipFree(dh);
    return e;
}
