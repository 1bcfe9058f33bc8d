// Launchers of the gfx950 stacking kernels (stack_kernel.hip), called by the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "stack_plan.hpp"

namespace ftk {

// One launch of the stacking kernel over all tiles of output rows [y_begin, y_end)
// and the block segments [seg0, seg0 + nseg) (bps blocks each). The frames buffer
// holds frame rows [row0, row1) of every frame, frame i at frames + i * frame_stride.
struct StackLaunch {
    const float* frames = nullptr;
    long long frame_stride = 0;
    int row0 = 0, row1 = 0;
    const int* shift = nullptr;
    const int* box = nullptr;
    const int* ubox = nullptr;
    const int* lds_off = nullptr;
    const long long* cov_off = nullptr;  // per block: offsets into cov_x, cov_y
    const int* cov_x = nullptr;
    const int* cov_y = nullptr;
    const int* motion = nullptr;
    const float* inv_sqrt = nullptr;
    int y_begin = 0, y_end = 0, tiles_x = 0, tiles_y = 0;
    int bps = 1, seg0 = 0, nseg = 1;
    unsigned long long* keys = nullptr;  // (y_end - y_begin) x w packed maxima
};

hipError_t launch_stack(const StackPlan& p, const StackLaunch& L, hipStream_t s);
hipError_t launch_fill_keys(unsigned long long* keys, long long count, hipStream_t s);
hipError_t launch_finalize(const unsigned long long* keys, long long count, float* max_snr, int* best_v,
                           hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, long long count, hipStream_t s);

}  // namespace ftk
