// 16-bit frame storage (fp16 / bf16 with a power-of-two scale per frame) and the
// FP64 evaluation of the stacking contract, for the precision form of the stacking
// pass.
//
// Storage: for frame i, s_i = 2^(e_i - 14) where max|g_i| = m 2^e_i (m in [0.5, 1));
// q = RNE16(g / s_i) (the division by a power of two is exact), so |q| < 2^14 and
// the decoded value q * s_i is exact in fp32. The stacking kernel sums the decoded
// values in fp32 in the contract's order, so a 16-bit pass equals the fp32 pass over
// the decoded frames bit for bit; the storage rounding (2^-11 relative for fp16, 2^-8
// for bf16) is the only error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/ftk_hip.h"
#include "stack_launch.hpp"
#include "store.hpp"

// Scratch of the allocation-free FP64 evaluation (ftk_stack_fp64_noalloc): the chunks'
// partial maxima of a strip of rows. Each GPU the library runs on holds its own copy
// for the life of the process. Host code never uses this name as an address (on the
// host it is a stand-in, not device memory): it takes the per-device address from
// hipGetSymbolAddress (scratch_base below).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace ftk {

namespace {

// maximum of |v| over the finite samples of frame blockIdx.y, as ordered bits (a
// non-negative float orders like its bit pattern)
__global__ void absmax_kernel(const float* __restrict__ src, long long stride, int pitch, int w, int h,
                              unsigned* __restrict__ maxbits) {
    const int i = blockIdx.y;
    const float* f = src + i * stride;
    const long long count = static_cast<long long>(w) * h;
    float m = 0.0f;
    for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < count;
         q += static_cast<long long>(gridDim.x) * blockDim.x) {
        const long long y = q / w;
        const float v = fabsf(f[y * pitch + (q - y * w)]);
        if (v <= 3.402823466e38f) m = fmaxf(m, v);  // NaN and inf do not set the range
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&maxbits[i], __float_as_uint(red[0]));
}

__global__ void scale_kernel(const unsigned* __restrict__ maxbits, int n, float* __restrict__ scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = __uint_as_float(maxbits[i]);
    int e = 0;
    if (m > 0.0f) (void)frexpf(m, &e);
    scale[i] = m > 0.0f ? ldexpf(1.0f, max(e - 14, -126)) : 1.0f;  // a normal number, so 1/scale is finite
}

__device__ inline uint16_t to_f16(float x) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x)); }

__device__ inline uint16_t to_bf16(float x) {
    const uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x0040u);  // quiet NaN
    return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);                  // nearest, ties even
}

__global__ void encode_kernel(const float* __restrict__ src, long long sstride, int spitch, uint16_t* __restrict__ dst,
                              long long dstride, int dpitch, const float* __restrict__ scale, int w, int storage) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const float inv = 1.0f / scale[i];  // a power of two: exact
    const float v = src[i * sstride + static_cast<long long>(y) * spitch + x] * inv;
    dst[i * dstride + static_cast<long long>(y) * dpitch + x] = storage == kStoreBF16 ? to_bf16(v) : to_f16(v);
}

// FP64 evaluation, phase 1: one thread per output pixel of rows [y0, y1) and one chunk
// of motions [j0, j0 + kFp64Chunk): for each motion j (ascending) the sum over covering
// frames i (ascending) in double, times 1/sqrt(c) in double; the strict maximum over
// the chunk (ties to the lowest index; a NaN sum is never selected) goes to the chunk's
// slot. Phase 2 takes the first maximum over the chunks in order, which is the strict
// maximum over all motions with ties to the lowest index.
constexpr int kFp64Chunk = 128;

__global__ void fp64_chunk_kernel(const float* __restrict__ frames, long long stride, int n, int w, int h, int nv,
                                  const int* __restrict__ dx, const int* __restrict__ dy, int min_count, int y0, int y1,
                                  int ybase, long long npx, double* __restrict__ cval, int* __restrict__ cidx) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = y0 + blockIdx.y;
    if (x >= w || y >= y1) return;
    const int j0 = blockIdx.z * kFp64Chunk, j1 = min(nv, j0 + kFp64Chunk);
    const __attribute__((address_space(4))) int* cdx = (const __attribute__((address_space(4))) int*)(dx);
    const __attribute__((address_space(4))) int* cdy = (const __attribute__((address_space(4))) int*)(dy);
    double bv = -HUGE_VAL;
    int bj = -1;
    for (int j = j0; j < j1; ++j) {
        double s = 0.0;
        int c = 0;
        for (int i = 0; i < n; ++i) {
            const int sx = x + cdx[static_cast<long long>(j) * n + i], sy = y + cdy[static_cast<long long>(j) * n + i];
            if (static_cast<unsigned>(sx) < static_cast<unsigned>(w) && static_cast<unsigned>(sy) < static_cast<unsigned>(h)) {
                s += static_cast<double>(frames[i * stride + static_cast<long long>(sy) * w + sx]);
                ++c;
            }
        }
        if (c < min_count) continue;
        const double v = s * (1.0 / sqrt(static_cast<double>(c)));
        if (v > bv) {
            bv = v;
            bj = j;
        }
    }
    const long long o = static_cast<long long>(blockIdx.z) * npx + static_cast<long long>(y - ybase) * w + x;
    cval[o] = bv;
    cidx[o] = bj;
}

__global__ void fp64_reduce_kernel(const double* __restrict__ cval, const int* __restrict__ cidx, int chunks,
                                   long long npx, double* __restrict__ out, int* __restrict__ best) {
    const long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= npx) return;
    double bv = -HUGE_VAL;
    int bj = -1;
    for (int c = 0; c < chunks; ++c) {
        const int j = cidx[c * npx + q];
        const double v = cval[c * npx + q];
        if (j >= 0 && v > bv) {
            bv = v;
            bj = j;
        }
    }
    out[q] = bj < 0 ? 0.0 : bv;
    best[q] = bj;
}

}  // namespace

hipError_t launch_store(const float* src, long long sstride, int spitch, int n, int w, int h, int storage,
                        uint16_t* dst, long long dstride, int dpitch, unsigned* maxbits, float* scale, hipStream_t s) {
    if (n <= 0 || w <= 0 || h <= 0) return hipSuccess;
    if (storage != kStoreF16 && storage != kStoreBF16) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(maxbits, 0, sizeof(unsigned) * n, s);
    if (e != hipSuccess) return e;
    const long long count = static_cast<long long>(w) * h;
    const unsigned bx = static_cast<unsigned>(std::min<long long>(1024, (count + 255) / 256));
    hipLaunchKernelGGL(absmax_kernel, dim3(bx, n), dim3(256), 0, s, src, sstride, spitch, w, h, maxbits);
    hipLaunchKernelGGL(scale_kernel, dim3((n + 63) / 64), dim3(64), 0, s, maxbits, n, scale);
    for (int y = 0; y < h; y += 32768) {
        const int ry = std::min(32768, h - y);
        hipLaunchKernelGGL(encode_kernel, dim3((w + 255) / 256, ry, n), dim3(256), 0, s,
                           src + static_cast<long long>(y) * spitch, sstride, spitch,
                           dst + static_cast<long long>(y) * dpitch, dstride, dpitch, scale, w, storage);
    }
    return hipGetLastError();
}

hipError_t launch_fp64(const float* frames, long long stride, int n, int w, int h, int nv, const int* d_dx,
                       const int* d_dy, int min_count, int y0, int y1, double* out, int* best, hipStream_t s) {
    if (y1 <= y0) return hipSuccess;
    const int chunks = (nv + kFp64Chunk - 1) / kFp64Chunk;
    const long long npx = static_cast<long long>(y1 - y0) * w;
    double* cval = nullptr;
    int* cidx = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&cval), sizeof(double) * chunks * npx, s);
    if (e != hipSuccess) return e;
    e = hipMallocAsync(reinterpret_cast<void**>(&cidx), sizeof(int) * chunks * npx, s);
    if (e != hipSuccess) {
        (void)hipFreeAsync(cval, s);
        return e;
    }
    for (int r = y0; r < y1; r += 32768) {
        const int r1 = std::min(y1, r + 32768);
        // rows [r, r1) write their slots of the [y0, y1) layout
        hipLaunchKernelGGL(fp64_chunk_kernel, dim3((w + 63) / 64, r1 - r, chunks), dim3(64), 0, s, frames, stride, n, w, h,
                           nv, d_dx, d_dy, min_count, r, r1, y0, npx, cval, cidx);
    }
    hipLaunchKernelGGL(fp64_reduce_kernel, dim3(static_cast<unsigned>((npx + 255) / 256)), dim3(256), 0, s, cval, cidx,
                       chunks, npx, out, best);
    e = hipGetLastError();
    (void)hipFreeAsync(cval, s);
    (void)hipFreeAsync(cidx, s);
    return e;
}

// The FP64 evaluation with its partials in `base` (bytes of device scratch): rows go
// in strips whose chunk partials (8 + 4 bytes per chunk and pixel) fit.
hipError_t launch_fp64_in(const float* frames, long long stride, int n, int w, int h, int nv, const int* d_dx,
                          const int* d_dy, int min_count, int y0, int y1, double* out, int* best, void* base,
                          size_t bytes, hipStream_t s) {
    if (y1 <= y0) return hipSuccess;
    const int chunks = (nv + kFp64Chunk - 1) / kFp64Chunk;
    const size_t row_bytes = static_cast<size_t>(w) * chunks * (sizeof(double) + sizeof(int));
    if (row_bytes > bytes) return hipErrorInvalidValue;
    const int rows = static_cast<int>(std::min<size_t>(32768, bytes / row_bytes));
    for (int r = y0; r < y1; r += rows) {
        const int r1 = std::min(y1, r + rows);
        const long long npx = static_cast<long long>(r1 - r) * w;
        double* cval = static_cast<double*>(base);
        int* cidx = reinterpret_cast<int*>(cval + static_cast<long long>(chunks) * npx);
        hipLaunchKernelGGL(fp64_chunk_kernel, dim3((w + 63) / 64, r1 - r, chunks), dim3(64), 0, s, frames, stride, n, w,
                           h, nv, d_dx, d_dy, min_count, r, r1, r, npx, cval, cidx);
        hipLaunchKernelGGL(fp64_reduce_kernel, dim3(static_cast<unsigned>((npx + 255) / 256)), dim3(256), 0, s, cval,
                           cidx, chunks, npx, out + static_cast<long long>(r - y0) * w,
                           best + static_cast<long long>(r - y0) * w);
    }
    return hipGetLastError();
}

namespace {

// Per-device address of `scratch`, taken once per device with that device current.
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_base[kMaxDevices] = {};

hipError_t scratch_base(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_base[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_base[dev] = p;
    }
    *out = g_scratch_base[dev];
    return hipSuccess;
}

void put_err3(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    const size_t k = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = '\0';
}

}  // namespace

}  // namespace ftk

extern "C" int ftk_stack_fp64_noalloc(int n, int w, int h, const float* d_frames, size_t frame_stride, int nv,
                                      const int32_t* d_dx, const int32_t* d_dy, int min_count, int y0, int y1,
                                      double* d_max_snr, int32_t* d_best_v, void* stream, char* err, size_t err_len) {
    using namespace ftk;
    std::string msg;
    if (!d_frames || !d_dx || !d_dy || !d_max_snr || !d_best_v) msg = "ftk_stack_fp64_noalloc: null argument";
    else if (n < 1 || w < 1 || h < 1 || nv < 1 || min_count < 1) msg = "ftk_stack_fp64_noalloc: bad sizes";
    else if (y0 < 0 || y1 > h || y1 < y0) msg = "ftk_stack_fp64_noalloc: row range outside the frame";
    else if (frame_stride < static_cast<size_t>(w) * h) msg = "ftk_stack_fp64_noalloc: frame stride smaller than a frame";
    else {
        const int chunks = (nv + kFp64Chunk - 1) / kFp64Chunk;
        const size_t row_bytes = static_cast<size_t>(w) * chunks * (sizeof(double) + sizeof(int));
        if (row_bytes > kNoAllocScratchBytes)
            msg = "ftk_stack_fp64_noalloc: one row of partial maxima needs " + std::to_string(row_bytes) +
                  " bytes, more than the " + std::to_string(kNoAllocScratchBytes) + "-byte scratch (w x ceil(nv / " +
                  std::to_string(kFp64Chunk) + ") x 12 bytes)";
    }
    if (!msg.empty()) {
        put_err3(err, err_len, msg);
        return FTK_ERROR;
    }
    void* base = nullptr;
    hipError_t e = scratch_base(&base);
    if (e != hipSuccess) {
        put_err3(err, err_len, std::string("ftk_stack_fp64_noalloc: scratch address: ") + hipGetErrorString(e));
        return FTK_ERROR;
    }
    e = launch_fp64_in(d_frames, static_cast<long long>(frame_stride), n, w, h, nv, d_dx, d_dy, min_count, y0, y1,
                       d_max_snr, d_best_v, base, kNoAllocScratchBytes, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        put_err3(err, err_len, std::string("ftk_stack_fp64_noalloc: ") + hipGetErrorString(e));
        return FTK_ERROR;
    }
    return FTK_OK;
}
