// gpv_probe.hip -- a kernel of known FP32 VALU instruction count, for the
// VALU roofline: its measured rate is the issue peak the Level2 kernels are
// compared with, and its SQ_INSTS_VALU counter reading checks that the
// counter counts wave instructions.  Built with the library's flags (packed
// FP32 off, no contraction), so every add below is one v_add_f32.
#include "gpv.h"
#include "gpv_internal.hpp"

#include <hip/hip_runtime.h>

namespace gpv {

constexpr int kProbeChains = 8;     // independent accumulators per thread
constexpr int kProbeUnroll = 4;     // loop body: kProbeChains * kProbeUnroll adds

__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float step)
{
    float x[kProbeChains];
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) x[k] = (float)(threadIdx.x + k);
    const float c = step * (float)(blockIdx.x & 7);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < kProbeUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < kProbeChains; ++k) x[k] = x[k] + c;
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kProbeChains; ++k) s = s + x[k];
    // a store the compiler cannot drop (step is never exactly this value)
    if (s == 1234.5f) out[blockIdx.x] = s;
}

}  // namespace gpv

extern "C" int gpv_valu_probe(gpv_ctx* ctx, int blocks, int iters, float* ms, double* wave_insts)
{
    (void)ctx;
    if (blocks < 1 || iters < 1 || !ms || !wave_insts) return gpv::set_error(GPV_ERR_INVALID, "bad probe arguments");
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return gpv::set_error(GPV_ERR_HIP, "event creation failed");
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return gpv::set_error(GPV_ERR_HIP, "event creation failed");
    }
    hipStream_t s = nullptr;
    int st = GPV_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) st = GPV_ERR_HIP;
    if (st == GPV_OK) {
        gpv::k_valu_probe<<<blocks, 256, 0, s>>>(nullptr, iters, 0.5f);   // warm-up
        (void)hipEventRecord(a, s);
        gpv::k_valu_probe<<<blocks, 256, 0, s>>>(nullptr, iters, 0.5f);
        (void)hipEventRecord(b, s);
        if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) st = GPV_ERR_HIP;
    }
    if (st == GPV_OK && hipEventElapsedTime(ms, a, b) != hipSuccess) st = GPV_ERR_HIP;
    // adds in the loop: 4 waves per block, iters * chains * unroll per wave
    *wave_insts = (double)blocks * 4.0 * (double)iters * gpv::kProbeChains * gpv::kProbeUnroll;
    if (s) (void)hipStreamDestroy(s);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return st == GPV_OK ? GPV_OK : gpv::set_error(GPV_ERR_HIP, "probe launch failed");
}
