// gfx950 kernels of the search loop around the stacking pass: local maxima of the
// max-S/N map (find_peaks, src/search.cpp:105-135) and the FP64 null-tail fit's
// normal equations (fit_null_tail_gauss, src/search.cpp:480-499).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "search_device.hpp"

namespace ftk {

// A pixel in [edge, w-edge) x [edge, h-edge) is a peak when its value is at least
// `floor`, no pixel of its (2r+1)^2 window is larger, and no earlier pixel (row-major)
// of the window is equal: one peak per plateau (search.cpp:112-129). Peaks are
// appended in any order; the host restores row-major order before its sort.
__global__ void peaks_kernel(const float* __restrict__ m, const int* __restrict__ best_v, int w, int h, float floor,
                             int radius, int edge, int cap, int* __restrict__ count, int* __restrict__ px,
                             int* __restrict__ py, float* __restrict__ ps, int* __restrict__ pv) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < edge || x >= w - edge || y < edge || y >= h - edge) return;
    const float v = m[static_cast<long long>(y) * w + x];
    if (v < floor) return;
    for (int dy = -radius; dy <= radius; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= h) continue;
        for (int dx = -radius; dx <= radius; ++dx) {
            const int xx = x + dx;
            if ((dx == 0 && dy == 0) || xx < 0 || xx >= w) continue;
            const float u = m[static_cast<long long>(yy) * w + xx];
            const bool earlier = dy < 0 || (dy == 0 && dx < 0);
            if (u > v || (u == v && earlier)) return;
        }
    }
    const int slot = atomicAdd(count, 1);
    if (slot < cap) {
        px[slot] = x;
        py[slot] = y;
        ps[slot] = v;
        pv[slot] = best_v[static_cast<long long>(y) * w + x];
    }
}

// Normal equations of log N(>u) = a + k log u - inv_s2 u^2 / 2 over the top m pooled
// peaks (sorted descending): sums of f f^T and f y with f = [1, log u, -u^2/2],
// y = log((j+1)/runs). One thread per peak, FP64, partial sums per workgroup.
__global__ void tail_sums_kernel(const float* __restrict__ pooled_desc, int m, double runs,
                                 double* __restrict__ partial, int* __restrict__ bad) {
    __shared__ double red[9][256];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (j < m) {
        const double u = pooled_desc[j];
        if (!(u > 0.0)) atomicOr(bad, 1);
        const double f0 = 1.0, f1 = log(u), f2 = -0.5 * u * u;
        const double yv = log(static_cast<double>(j + 1) / runs);
        s[0] = f0 * f0; s[1] = f0 * f1; s[2] = f0 * f2;
        s[3] = f1 * f1; s[4] = f1 * f2; s[5] = f2 * f2;
        s[6] = f0 * yv; s[7] = f1 * yv; s[8] = f2 * yv;
    }
    for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 9; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 9; ++q) partial[blockIdx.x * 9 + q] = red[q][0];
}

// Per-frame rows of a stacking plan's tables for a reordered frame sequence:
// dst[(b n + i) row + e] = src[(b n + perm[i]) row + e]. A null run that assigns the
// real pass's frame times in the order perm stacks with exactly the real pass's shifts,
// frame i taking the shifts of frame perm[i], so its plan is the real plan with these
// rows permuted (the blocks, coverage tables and bounds depend only on the set of
// shifts per motion, which the permutation does not change).
// Up to four per-frame tables permuted in one launch: table t has nb[t] blocks of n
// frame rows of row[t] ints; row i of block b of dst is row perm[i] of block b of src.
// 32-bit indices (every plan table here is far below 2^31 ints).
struct PermuteTables {
    const int* src[4];
    int* dst[4];
    int nb[4], row[4];
    int total[4];  // nb * n * row
    int count;
};

__global__ void permute_rows_kernel(PermuteTables t, int n, const int* __restrict__ perm) {
    const int stride = gridDim.x * blockDim.x;
    for (int k = 0; k < t.count; ++k) {
        const int row = t.row[k];
        const int* __restrict__ src = t.src[k];
        int* __restrict__ dst = t.dst[k];
        for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < t.total[k]; q += stride) {
            const int r = q / row, e = q - r * row;
            const int b = r / n, i = r - b * n;
            dst[q] = src[(b * n + perm[i]) * row + e];
        }
    }
}

hipError_t launch_permute_rows(const PermuteTable* tables, int count, int n, const int* perm, hipStream_t s) {
    PermuteTables t{};
    long long most = 0;
    t.count = 0;
    for (int k = 0; k < count && k < 4; ++k) {
        const long long total = static_cast<long long>(tables[k].nb) * n * tables[k].row;
        if (total <= 0) continue;
        if (total > 0x7fffffffLL) return hipErrorInvalidValue;
        const int j = t.count++;
        t.src[j] = tables[k].src;
        t.dst[j] = tables[k].dst;
        t.nb[j] = tables[k].nb;
        t.row[j] = tables[k].row;
        t.total[j] = static_cast<int>(total);
        most = std::max(most, total);
    }
    if (t.count == 0) return hipSuccess;
    const unsigned blocks = static_cast<unsigned>(std::min<long long>((most + 255) / 256, 4096));
    hipLaunchKernelGGL(permute_rows_kernel, dim3(blocks), dim3(256), 0, s, t, n, perm);
    return hipGetLastError();
}

hipError_t launch_peaks(const float* m, const int* best_v, int w, int h, float floor, int radius, int edge, int cap,
                        int* count, int* px, int* py, float* ps, int* pv, hipStream_t s) {
    hipLaunchKernelGGL(peaks_kernel, dim3((w + 127) / 128, h), dim3(128), 0, s, m, best_v, w, h, floor, radius, edge,
                       cap, count, px, py, ps, pv);
    return hipGetLastError();
}

hipError_t launch_tail_sums(const float* pooled_desc, int m, double runs, double* partial, int* bad, hipStream_t s) {
    const int blocks = (m + 255) / 256;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(tail_sums_kernel, dim3(blocks), dim3(256), 0, s, pooled_desc, m, runs, partial, bad);
    return hipGetLastError();
}

}  // namespace ftk
