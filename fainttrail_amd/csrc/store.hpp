// Launchers of store_kernels.hip: 16-bit frame storage and the FP64 evaluation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ftk {

// n frames of w x h fp32 (rows spitch floats apart, frames sstride apart) to 16-bit
// storage (kStoreF16 / kStoreBF16) at dst (rows dpitch, frames dstride elements apart)
// with the power-of-two scale of each frame written to scale[n]; maxbits: n words of
// device scratch.
hipError_t launch_store(const float* src, long long sstride, int spitch, int n, int w, int h, int storage,
                        uint16_t* dst, long long dstride, int dpitch, unsigned* maxbits, float* scale, hipStream_t s);

// The stacking contract in FP64 for output rows [y0, y1): out / best hold (y1 - y0) x w
// values. dx, dy: device shift tables, motion-major (j n + i).
hipError_t launch_fp64(const float* frames, long long stride, int n, int w, int h, int nv, const int* d_dx,
                       const int* d_dy, int min_count, int y0, int y1, double* out, int* best, hipStream_t s);

}  // namespace ftk
