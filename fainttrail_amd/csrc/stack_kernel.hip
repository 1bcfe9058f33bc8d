// gfx950 shift-and-stack kernel with per-pixel max S/N and argmax.
//
// Contract (include/fainttrail/backend.hpp:10-40): for output pixel (x, y) and trial
// motion j, S_j = (sum over covering frames i, ascending, of g_i(x+dx_ji, y+dy_ji))
// * float(1/sqrt(c)), summed in fp32 from +0.0f; undefined when c < min_count;
// max over j with ties to the lowest j. Bit-identical to the CPU backend
// (src/backend_cpu.cpp:77-204) because
//   * every (pixel, motion) sum adds the covered samples in ascending frame order,
//     and an uncovered frame adds -0.0f, the exact identity of IEEE addition
//     (a sum that starts at +0.0f is never -0.0f, so +0 + -0 = +0 keeps it);
//   * the build uses -ffp-contract=off, so no fused multiply-add appears;
//   * c and 1/sqrt(c) come from the same integer count and the same host table;
//   * the maximum keeps (value, lowest index), merged across workgroups with a
//     64-bit atomic max of (ordered value bits, inverted index).
//
// One workgroup = one tile (64 x 16 pixels; each wavefront owns 4 rows of 64) and a
// segment of velocity blocks. Per block and frame, the box the block's shifted tiles
// can reach is staged in LDS (double buffered, one barrier per frame) and each lane
// reads its 16 motions x 4 rows from it. No MFMA: this is a gather-sum.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "stack_launch.hpp"
#include "stack_plan.hpp"

namespace ftk {

namespace {

constexpr int V = kVel;
constexpr int RPW = kRowsPerWave;
constexpr int kPreRows = kLdsRows / kWaves;  // staged rows per wavefront, at most
constexpr int kPreCols = (kPitch + kTileX - 1) / kTileX;

__device__ inline uint32_t d_ordered(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float d_unordered(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

}  // namespace

__global__ __launch_bounds__(kThreads) void stack_kernel(
    const float* __restrict__ frames, long long frame_stride, int row0, int row1, int n, int w, int h,
    const int* __restrict__ shift, const int* __restrict__ box, const int* __restrict__ ubox,
    const int* __restrict__ motion, const float* __restrict__ inv_sqrt, int min_count, int nb, int bps,
    int tiles_x, int y_begin, int y_end, int seg0, int nseg, unsigned long long* __restrict__ keys) {
    __shared__ float lds[2][kLdsRows * kPitch];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // Workgroups go to the 8 XCDs round-robin. Renumber so that each XCD works through
    // a contiguous range of (tile, segment) pairs: all segments of a tile then run on
    // one XCD and share its L2 copy of the tile's frame boxes.
    const unsigned total = gridDim.x, b = blockIdx.x;
    const unsigned q = total >> 3, r8 = total & 7u, xcd = b & 7u;
    const unsigned logical = xcd * q + (xcd < r8 ? xcd : r8) + (b >> 3);
    const int tile = static_cast<int>(logical / static_cast<unsigned>(nseg));
    const int seg = seg0 + static_cast<int>(logical % static_cast<unsigned>(nseg));
    const int x0 = (tile % tiles_x) * kTileX;
    const int y0 = y_begin + (tile / tiles_x) * kTileY;
    const int x = x0 + lane;
    const int yw = y0 + wave * RPW;
    const float inv_n = inv_sqrt[n];

    float best[RPW];
    int arg[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        best[r] = -__builtin_inff();
        arg[r] = -1;
    }

    const int blk_begin = seg * bps;
    const int blk_end = min(nb, blk_begin + bps);
    for (int blk = blk_begin; blk < blk_end; ++blk) {
        const int* ub = ubox + 4 * blk;
        // interior: every staged read of every frame lies in the frame rows present
        const bool interior = x0 + ub[0] >= 0 && x0 + kTileX - 1 + ub[2] < w && y0 + ub[1] >= row0 &&
                              y0 + kTileY - 1 + ub[3] < row1;

        float acc[V][RPW];
#pragma unroll
        for (int k = 0; k < V; ++k)
#pragma unroll
            for (int r = 0; r < RPW; ++r) acc[k][r] = 0.0f;

        float pre[kPreRows][kPreCols];
        // gather frame i's box into registers (coalesced rows; -0.0f outside the frame)
        auto gather = [&](int i) {
            const int* bx = box + 4 * (static_cast<long long>(blk) * n + i);
            const int bx0 = x0 + bx[0], by0 = y0 + bx[1];
            const int C = kTileX + bx[2], R = kTileY + bx[3];
            const float* f = frames + static_cast<long long>(i) * frame_stride;
#pragma unroll
            for (int m = 0; m < kPreRows; ++m) {
                const int rr = wave + kWaves * m;
                const int gy = by0 + rr;
#pragma unroll
                for (int c = 0; c < kPreCols; ++c) {
                    const int cc = lane + kTileX * c;
                    const int gx = bx0 + cc;
                    float v = -0.0f;
                    if (rr < R && cc < C) {
                        if (interior) {
                            v = f[static_cast<long long>(gy - row0) * w + gx];
                        } else if (gy >= row0 && gy < row1 && gx >= 0 && gx < w) {
                            v = f[static_cast<long long>(gy - row0) * w + gx];
                        }
                    }
                    pre[m][c] = v;
                }
            }
        };
        auto deposit = [&](int buf, int i) {
            const int* bx = box + 4 * (static_cast<long long>(blk) * n + i);
            const int C = kTileX + bx[2], R = kTileY + bx[3];
#pragma unroll
            for (int m = 0; m < kPreRows; ++m) {
                const int rr = wave + kWaves * m;
#pragma unroll
                for (int c = 0; c < kPreCols; ++c) {
                    const int cc = lane + kTileX * c;
                    if (rr < R && cc < C) lds[buf][rr * kPitch + cc] = pre[m][c];
                }
            }
        };

        gather(0);
        deposit(0, 0);
        __syncthreads();
        for (int i = 0; i < n; ++i) {
            const int cur = i & 1;
            if (i + 1 < n) gather(i + 1);
            const int* bx = box + 4 * (static_cast<long long>(blk) * n + i);
            const int mdx = bx[0], mdy = bx[1];
            const int* sh = shift + (static_cast<long long>(blk) * n + i) * (2 * V);
            const float* L = lds[cur];
            const int lane_base = wave * RPW * kPitch + lane;
#pragma unroll
            for (int k = 0; k < V; ++k) {
                const int off = lane_base + (sh[V + k] - mdy) * kPitch + (sh[k] - mdx);
#pragma unroll
                for (int r = 0; r < RPW; ++r) acc[k][r] += L[off + r * kPitch];
            }
            if (i + 1 < n) deposit(cur ^ 1, i + 1);
            __syncthreads();
        }

        // S/N of the block's motions (ascending index) and the running maximum
        const int* mo = motion + V * blk;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const int j = mo[k];
            if (j < 0) continue;
            int cnt[RPW];
            if (interior) {
#pragma unroll
                for (int r = 0; r < RPW; ++r) cnt[r] = n;
            } else {
                // coverage from the full-frame bounds, frame by frame
#pragma unroll
                for (int r = 0; r < RPW; ++r) cnt[r] = 0;
                for (int i = 0; i < n; ++i) {
                    const int* sh = shift + (static_cast<long long>(blk) * n + i) * (2 * V);
                    const int sdx = sh[k], sdy = sh[V + k];
                    const int xin = static_cast<unsigned>(x + sdx) < static_cast<unsigned>(w) ? 1 : 0;
#pragma unroll
                    for (int r = 0; r < RPW; ++r)
                        cnt[r] += static_cast<unsigned>(yw + r + sdy) < static_cast<unsigned>(h) ? xin : 0;
                }
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                if (cnt[r] < min_count) continue;
                const float snr = acc[k][r] * (interior ? inv_n : inv_sqrt[cnt[r]]);
                if (snr > best[r] || (snr == best[r] && j < arg[r])) {
                    best[r] = snr;
                    arg[r] = j;
                }
            }
        }
    }

#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int y = yw + r;
        if (arg[r] < 0 || x >= w || y >= y_end) continue;
        const unsigned long long key =
            (static_cast<unsigned long long>(d_ordered(best[r])) << 32) | static_cast<uint32_t>(~static_cast<uint32_t>(arg[r]));
        atomicMax(&keys[static_cast<long long>(y - y_begin) * w + x], key);
    }
}

__global__ void fill_keys_kernel(unsigned long long* __restrict__ keys, long long count) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) keys[i] = 0ull;
}

// Keys -> (max_snr, best_v): 0 / -1 where no motion is defined; a non-finite maximum
// becomes 0 with its motion kept (src/backend_cpu.cpp:191-192).
__global__ void finalize_kernel(const unsigned long long* __restrict__ keys, long long count,
                                float* __restrict__ max_snr, int* __restrict__ best_v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const unsigned long long k = keys[i];
    if (k == 0ull) {
        max_snr[i] = 0.0f;
        best_v[i] = -1;
        return;
    }
    const float v = d_unordered(static_cast<uint32_t>(k >> 32));
    max_snr[i] = __builtin_isfinite(v) ? v : 0.0f;
    best_v[i] = static_cast<int>(~static_cast<uint32_t>(k & 0xffffffffull));
}

__global__ void copy_f32_kernel(const float* __restrict__ src, float* __restrict__ dst, long long count) {
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
        dst[i] = src[i];
}

// ------------------------------------------------------------------ launchers

hipError_t launch_fill_keys(unsigned long long* keys, long long count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(fill_keys_kernel, dim3(blocks), dim3(256), 0, s, keys, count);
    return hipGetLastError();
}

hipError_t launch_finalize(const unsigned long long* keys, long long count, float* max_snr, int* best_v,
                           hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(256), 0, s, keys, count, max_snr, best_v);
    return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, long long count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_f32_kernel, dim3(2048), dim3(256), 0, s, src, dst, count);
    return hipGetLastError();
}

hipError_t launch_stack(const StackPlan& p, const StackLaunch& L, hipStream_t s) {
    const unsigned long long wgs = static_cast<unsigned long long>(L.tiles_x) * L.tiles_y * L.nseg;
    if (wgs == 0) return hipSuccess;
    hipLaunchKernelGGL(stack_kernel, dim3(static_cast<unsigned>(wgs)), dim3(kThreads), 0, s, L.frames,
                       L.frame_stride, L.row0, L.row1, p.n, p.w, p.h, L.shift, L.box, L.ubox, L.motion, L.inv_sqrt,
                       p.min_count, p.nb, L.bps, L.tiles_x, L.y_begin, L.y_end, L.seg0, L.nseg, L.keys);
    return hipGetLastError();
}

}  // namespace ftk
