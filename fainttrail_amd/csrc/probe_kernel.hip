// LDS read-bandwidth probe: the ceiling of the stacking kernel, whose inner loop is
// one LDS read and one fp32 add per stacked unit. Each workgroup fills 32 KiB of LDS
// and streams wave-contiguous reads of `width` bytes per lane through it, summing what
// it reads so no read is dead. Returns achieved bytes per second. Widths: 4
// (ds_read_b32), 8 (two dwords 128 floats apart in one ds_read2st64_b32, the form of
// the fp32 stacking loop, which reads a lane's rows r and r + 1 of a 128-float pitch)
// and 16 (ds_read_b128).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/ftk_hip.h"

namespace ftk {

namespace {

constexpr int kProbeFloats = 8192;  // 32 KiB

template <int W>
__global__ __launch_bounds__(256) void lds_probe_kernel(float* __restrict__ out, int iters) {
    __shared__ float buf[kProbeFloats];
    for (int i = threadIdx.x; i < kProbeFloats; i += 256) buf[i] = static_cast<float>(i & 255) * 0.001f;
    __syncthreads();
    float acc[4 * W] = {};
    int base = threadIdx.x * W;
    const int lane = threadIdx.x & 63;
    int wbase = static_cast<int>(threadIdx.x >> 6) * 1024;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float* p = &buf[(base + k * 256 * W) & (kProbeFloats - 1)];
            if constexpr (W == 1) {
                acc[k] += p[0];
            } else if constexpr (W == 2) {
                // lane-consecutive dwords at q and q + 128 (q < 4096 + 64, in bounds)
                const float* r = &buf[((wbase + k * 256) & (kProbeFloats / 2 - 1)) + lane];
                acc[2 * k] += r[0];
                acc[2 * k + 1] += r[128];
            } else {
                const float4 v = *reinterpret_cast<const float4*>(p);
                acc[4 * k] += v.x;
                acc[4 * k + 1] += v.y;
                acc[4 * k + 2] += v.z;
                acc[4 * k + 3] += v.w;
            }
        }
        base = (base + 64 * W) & (kProbeFloats - 1);
        wbase = (wbase + 64) & (kProbeFloats / 2 - 1);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 4 * W; ++k) s += acc[k];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

}  // namespace ftk

extern "C" int ftk_probe_lds(int width_bytes, double* bytes_per_s) {
    using namespace ftk;
    if (!bytes_per_s || (width_bytes != 4 && width_bytes != 8 && width_bytes != 16)) return FTK_ERROR;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FTK_ERROR;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return FTK_ERROR;
    // five 32 KiB workgroups fill a CU's 160 KiB of LDS: one resident round, no tail;
    // the best of five timed launches of a few milliseconds each
    const int blocks = cus * 5, iters = 16384;
    float* out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * 256) != hipSuccess) return FTK_ERROR;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    auto launch = [&]() {
        if (width_bytes == 4) hipLaunchKernelGGL(lds_probe_kernel<1>, dim3(blocks), dim3(256), 0, nullptr, out, iters);
        else if (width_bytes == 8) hipLaunchKernelGGL(lds_probe_kernel<2>, dim3(blocks), dim3(256), 0, nullptr, out, iters);
        else hipLaunchKernelGGL(lds_probe_kernel<4>, dim3(blocks), dim3(256), 0, nullptr, out, iters);
    };
    launch();  // warm-up
    hipError_t e = hipSuccess;
    float ms = 1e30f;
    for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) {
        (void)hipEventRecord(a, nullptr);
        launch();
        (void)hipEventRecord(b, nullptr);
        e = hipEventSynchronize(b);
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, a, b);
        ms = std::min(ms, t);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(out);
    if (e != hipSuccess || ms <= 0.0f) return FTK_ERROR;
    const double bytes = static_cast<double>(blocks) * 256 * iters * 4 * width_bytes;
    *bytes_per_s = bytes / (ms * 1e-3);
    return FTK_OK;
}
