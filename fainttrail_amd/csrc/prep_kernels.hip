// gfx950 kernels of the time-order-dependent preprocessing stage (build_stage,
// src/preprocess.cpp:286-486) and the per-frame background (:528-585), bit-identical
// to the reference's CPU arithmetic:
//   * every float/double conversion sits where the reference has it;
//   * order statistics are exact whatever the algorithm (per-pixel rank counting in
//     LDS, per-frame radix select on ordered float bits), so any exact selection
//     reproduces std::sort / std::nth_element results;
//   * transcendental constants (filter taps, the cosmic-ray ratio, 1/sigma) come from
//     the host (glibc); device code uses only IEEE + - * / and comparisons, built with
//     -ffp-contract=off and correctly rounded fp32 division.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "prep.hpp"

namespace ftk {

namespace {

__device__ inline uint32_t ord_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float unord_bits(uint32_t o) { return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o); }
__device__ inline bool finitef(float v) { return __builtin_isfinite(v); }

}  // namespace

// ------------------------------------------------------------------ static sky
// Per pixel: the finite samples in (value, time rank) order (ranks are distinct, so the
// order is total, as std::sort of the pairs makes it, preprocess.cpp:171). Global
// median: the two middle values added in float, halved in double. Leave-out median for
// frame i: the sorted samples without the frames within `exclude` time ranks of i
// (preprocess.cpp:178-199), a and b in double, st = 0.5 (a + b); fewer than 5 left:
// the global median. Residual: float(double(v) - st).
//
// sky_rank_kernel (n <= NMAX <= 64): one pixel per lane, its samples in registers in
// TIME-RANK order (rv[r] = frame fr[r]'s sample), so the (value, rank) order needs no
// rank lookups: sample r's sorted position is the count of samples s < r with
// rv[s] <= rv[r] plus those s > r with rv[s] < rv[r] (non-finite samples are NaN and
// never counted). The sorted samples go to an LDS column (sv), each rank's sorted
// position to a byte column (pb). Frames are then visited in time-rank order with a
// 64-bit mask W of the sorted positions of the ranks within `exclude` of the current
// one, updated by one bit in and one out per step: the kept samples are the set bits of
// valid & ~W, their count a popcount, and the lo-th / hi-th kept sample the position
// of the lo-th set bit (a select by halving popcounts) and of the next set bit. Two LDS
// reads per frame instead of a walk over the sorted order.
namespace {
// f(integral_constant<int, I>) for I = B .. E - 1: the indices are constants before any
// loop pass runs, so register arrays indexed with them are split into registers at once.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

__device__ inline int select_bit64(uint64_t k_mask, int k) {
    // position of the k-th (from 0) set bit; the caller guarantees k < popcount
    uint32_t x = static_cast<uint32_t>(k_mask);
    int c = __builtin_popcount(x), pos = 0;
    const bool up = k >= c;
    k -= up ? c : 0;
    pos = up ? 32 : 0;
    x = up ? static_cast<uint32_t>(k_mask >> 32) : x;
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        const uint32_t lowmask = (1u << half) - 1u;
        c = __builtin_popcount(x & lowmask);
        const bool u = k >= c;
        k -= u ? c : 0;
        pos += u ? half : 0;
        x = u ? (x >> half) : x;
    }
    return pos;
}
}  // namespace

template <int NMAX>
__global__ __launch_bounds__(64) void sky_rank_kernel(const float* __restrict__ sub, long long fstride, int n,
                                                      long long plane, const int* __restrict__ rank, int exclude,
                                                      int min_finite, float* __restrict__ sky, float* __restrict__ resid,
                                                      long long rstride) {
    static_assert(NMAX >= 1 && NMAX <= 64, "one 64-bit mask of sorted positions");
    extern __shared__ unsigned char smem[];
    int* fr = reinterpret_cast<int*>(smem);                                   // [n]: frame of time rank r
    float* sv = reinterpret_cast<float*>(smem + 4 * n);                       // [n][64]: sorted samples
    unsigned char* pb = smem + 4 * n + 4 * 64 * n;                            // [n][64]: sorted position of rank r
    const int t = threadIdx.x;
    for (int i = t; i < n; i += 64) fr[rank[i]] = i;
    __syncthreads();  // the only barrier: lanes past the plane may leave below
    const long long p = static_cast<long long>(blockIdx.x) * 64 + t;
    if (p >= plane) return;
    float rv[NMAX];
    int m = 0;
    static_for<0, NMAX>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        float v = __builtin_nanf("");
        if (r < n) v = sub[fr[r] * fstride + p];
        const bool f = r < n && finitef(v);
        rv[r] = f ? v : __builtin_nanf("");
        m += f ? 1 : 0;
    });
    if (m < min_finite) {
        sky[p] = __builtin_nanf("");
        for (int i = 0; i < n; ++i) resid[i * rstride + p] = __builtin_nanf("");
        return;
    }
    static_for<0, NMAX>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        const float vr = rv[r];
        int pos = 0;
        static_for<0, NMAX>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q < r) pos += rv[q] <= vr ? 1 : 0;
            else if constexpr (q > r) pos += rv[q] < vr ? 1 : 0;
        });
        const bool f = vr == vr;
        if (r < n) pb[r * 64 + t] = f ? static_cast<unsigned char>(pos) : 255u;
        if (f) sv[pos * 64 + t] = vr;
    });
    const float ga = sv[((m - 1) / 2) * 64 + t];
    const float gb = sv[(m / 2) * 64 + t];
    const float gsum = ga + gb;
    const double global = 0.5 * static_cast<double>(gsum);
    sky[p] = static_cast<float>(global);
    const uint64_t valid = m >= 64 ? ~0ull : ((1ull << m) - 1ull);
    auto bit_of = [&](int r) -> uint64_t {
        const unsigned q = pb[r * 64 + t];
        return q < 64u ? (1ull << q) : 0ull;
    };
    // W: sorted positions of the ranks in [r - exclude, r + exclude] (frame r included)
    uint64_t W = 0ull;
    if (exclude > 0)
        for (int q = 0; q <= exclude && q < n; ++q) W |= bit_of(q);
    static_for<0, NMAX>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if (r >= n) return;
        if (r > 0 && exclude > 0) {
            if (r + exclude < n) W |= bit_of(r + exclude);
            if (r - exclude - 1 >= 0) W &= ~bit_of(r - exclude - 1);
        }
        const float v = rv[r];
        float out = __builtin_nanf("");
        if (v == v) {
            double st = global;
            if (exclude > 0) {
                const uint64_t K = valid & ~W;
                const int kept = __builtin_popcountll(K);
                if (kept >= 5) {
                    const int lo = (kept - 1) / 2, hi = kept / 2;
                    const int qlo = select_bit64(K, lo);
                    // hi is lo or lo + 1: the next set bit after qlo (qlo < 63 then)
                    const int qhi = hi == lo ? qlo : __builtin_ctzll(K & ~((2ull << qlo) - 1ull));
                    const double a = sv[qlo * 64 + t];
                    const double b = sv[qhi * 64 + t];
                    st = 0.5 * (a + b);
                }
            }
            out = static_cast<float>(static_cast<double>(v) - st);
        }
        resid[fr[r] * rstride + p] = out;
    });
}

// sky_kernel (any n): the pixel's n samples in an LDS column (vals[i][lane]); each
// finite sample's position in the (value, time rank) order is counted, giving the
// sorted index -> frame map `ord`; the leave-out median walks that order skipping the
// frames within `exclude` ranks of i.
__global__ __launch_bounds__(64) void sky_kernel(const float* __restrict__ sub, long long fstride, int n,
                                                 long long plane, const int* __restrict__ rank, int exclude,
                                                 int min_finite, float* __restrict__ sky, float* __restrict__ resid,
                                                 long long rstride) {
    extern __shared__ unsigned char smem[];
    float* vals = reinterpret_cast<float*>(smem);                        // [n][64]
    int* rk = reinterpret_cast<int*>(smem + 4 * 64 * n);                 // [n]
    uint16_t* ord = reinterpret_cast<uint16_t*>(smem + 4 * 64 * n + 4 * n);  // [n][64]
    const int t = threadIdx.x;
    for (int i = t; i < n; i += 64) rk[i] = rank[i];
    __syncthreads();
    const long long p = static_cast<long long>(blockIdx.x) * 64 + t;
    if (p >= plane) return;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const float v = sub[i * fstride + p];
        vals[i * 64 + t] = v;
        m += finitef(v) ? 1 : 0;
    }
    if (m < min_finite) {
        sky[p] = __builtin_nanf("");
        for (int i = 0; i < n; ++i) resid[i * rstride + p] = __builtin_nanf("");
        return;
    }
    for (int i = 0; i < n; ++i) {
        const float vi = vals[i * 64 + t];
        if (!finitef(vi)) continue;
        const int ri = rk[i];
        int pos = 0;
        for (int j = 0; j < n; ++j) {
            const float vj = vals[j * 64 + t];
            pos += (finitef(vj) && (vj < vi || (vj == vi && rk[j] < ri))) ? 1 : 0;
        }
        ord[pos * 64 + t] = static_cast<uint16_t>(i);
    }
    const float ga = vals[ord[((m - 1) / 2) * 64 + t] * 64 + t];
    const float gb = vals[ord[(m / 2) * 64 + t] * 64 + t];
    const float gsum = ga + gb;
    const double global = 0.5 * static_cast<double>(gsum);
    sky[p] = static_cast<float>(global);
    for (int i = 0; i < n; ++i) {
        const float v = vals[i * 64 + t];
        if (!finitef(v)) {
            resid[i * rstride + p] = __builtin_nanf("");
            continue;
        }
        double st = global;
        if (exclude > 0) {
            const int ri = rk[i];
            int kept = 0;
            for (int k = 0; k < m; ++k) {
                const int d = rk[ord[k * 64 + t]] - ri;
                kept += (d > exclude || d < -exclude) ? 1 : 0;
            }
            if (kept >= 5) {
                const int lo = (kept - 1) / 2, hi = kept / 2;
                double a = 0.0, b = 0.0;
                int c = 0;
                for (int k = 0; k < m; ++k) {
                    const int f = ord[k * 64 + t];
                    const int d = rk[f] - ri;
                    if (d <= exclude && d >= -exclude) continue;
                    if (c == lo) a = vals[f * 64 + t];
                    if (c == hi) {
                        b = vals[f * 64 + t];
                        break;
                    }
                    ++c;
                }
                st = 0.5 * (a + b);
            }
        }
        resid[i * rstride + p] = static_cast<float>(static_cast<double>(v) - st);
    }
}

// ------------------------------------------------------------------ scatter map
// scatter = 1.4826 * median_i(float(|r_i| / sigma_i)) over finite residuals, 1e9
// when fewer than n/2 are finite (preprocess.cpp:371-391). The median's order
// statistics come from rank counting in an LDS column.
__global__ __launch_bounds__(64) void scatter_kernel(const float* __restrict__ resid, long long rstride, int n,
                                                     long long plane, const double* __restrict__ sigma,
                                                     float* __restrict__ scatter) {
    extern __shared__ unsigned char smem[];
    float* col = reinterpret_cast<float*>(smem);  // [n][64]
    const int t = threadIdx.x;
    const long long p = static_cast<long long>(blockIdx.x) * 64 + t;
    if (p >= plane) return;
    int c = 0;
    for (int i = 0; i < n; ++i) {
        const float v = resid[i * rstride + p];
        if (!finitef(v)) continue;
        col[c * 64 + t] = static_cast<float>(static_cast<double>(__builtin_fabsf(v)) / sigma[i]);
        ++c;
    }
    if (c < n / 2) {
        scatter[p] = 1e9f;
        return;
    }
    const int mid = c / 2;
    float vmid = 0.0f, vlow = 0.0f;
    for (int j = 0; j < c; ++j) {
        const float vj = col[j * 64 + t];
        int lt = 0, le = 0;
        for (int k = 0; k < c; ++k) {
            const float vk = col[k * 64 + t];
            lt += vk < vj ? 1 : 0;
            le += vk <= vj ? 1 : 0;
        }
        if (lt <= mid && mid < le) vmid = vj;
        if (lt <= mid - 1 && mid - 1 < le) vlow = vj;
    }
    double med = vmid;
    if (c % 2 == 0) med = 0.5 * (med + static_cast<double>(vlow));
    scatter[p] = static_cast<float>(1.4826 * med);
}

// A bitonic sorting network over a register array (compile-time indices only, so the array
// stays in registers): ascending order by min / max, exact for values without NaN.
template <int N, int K, int J>
__device__ __forceinline__ void bitonic_pass(float (&u)[N]) {
    static_for<0, N>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        constexpr int l = i ^ J;
        if constexpr (l > i) {
            const float a = u[i], b = u[l];
            if constexpr ((i & K) == 0) {
                u[i] = __builtin_fminf(a, b);
                u[l] = __builtin_fmaxf(a, b);
            } else {
                u[i] = __builtin_fmaxf(a, b);
                u[l] = __builtin_fminf(a, b);
            }
        }
    });
}
template <int N, int K, int J>
__device__ __forceinline__ void bitonic_merge(float (&u)[N]) {
    if constexpr (J > 0) {
        bitonic_pass<N, K, J>(u);
        bitonic_merge<N, K, J / 2>(u);
    }
}
template <int N, int K = 2>
__device__ __forceinline__ void bitonic_sort(float (&u)[N]) {
    if constexpr (K <= N) {
        bitonic_merge<N, K, K / 2>(u);
        bitonic_sort<N, K * 2>(u);
    }
}

// scatter_kernel for n <= NMAX (a power of two, at most 64): the pixel's values
// float(|r_i| / sigma_i) in registers, +inf in the slots of non-finite or missing samples
// (the finite values can be +inf only by overflow, and then sort with them to the top of
// the c valid ones), sorted by the network; the middle elements are then picked with
// compile-time indices. Same values and the same double arithmetic as scatter_kernel.
template <int NMAX>
__global__ __launch_bounds__(64) void scatter_sort_kernel(const float* __restrict__ resid, long long rstride, int n,
                                                          long long plane, const double* __restrict__ sigma,
                                                          float* __restrict__ scatter) {
    const long long p = static_cast<long long>(blockIdx.x) * 64 + threadIdx.x;
    if (p >= plane) return;
    float u[NMAX];
    int c = 0;
    static_for<0, NMAX>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        float x = __builtin_inff();
        if (i < n) {
            const float v = resid[i * rstride + p];
            if (finitef(v)) {
                x = static_cast<float>(static_cast<double>(__builtin_fabsf(v)) / sigma[i]);
                ++c;
            }
        }
        u[i] = x;
    });
    if (c < n / 2) {
        scatter[p] = 1e9f;
        return;
    }
    bitonic_sort(u);
    const int mid = c / 2;
    float vmid = 0.0f, vlow = 0.0f;
    static_for<0, NMAX>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        vmid = k == mid ? u[k] : vmid;
        vlow = k == mid - 1 ? u[k] : vlow;
    });
    double med = vmid;
    if (c % 2 == 0) med = 0.5 * (med + static_cast<double>(vlow));
    scatter[p] = static_cast<float>(1.4826 * med);
}

// mask = 0 within dil pixels (square) of any pixel whose scatter exceeds mask_scatter
// (preprocess.cpp:393-400; the comparison is float against double).
__global__ void mask_kernel(const float* __restrict__ scatter, int w, int h, int dil, double mask_scatter,
                            float* __restrict__ mask) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    float out = 1.0f;
    for (int yy = max(0, y - dil); yy <= min(h - 1, y + dil) && out != 0.0f; ++yy)
        for (int xx = max(0, x - dil); xx <= min(w - 1, x + dil); ++xx)
            if (!(static_cast<double>(scatter[static_cast<long long>(yy) * w + xx]) <= mask_scatter)) {
                out = 0.0f;
                break;
            }
    mask[static_cast<long long>(y) * w + x] = out;
}

// z = r * float(1/sigma_i) / e where unmasked and finite, else 0; e = scatter when it
// exceeds weight_scatter, else 1 (preprocess.cpp:413-425). In place over r.
__global__ void z_kernel(float* __restrict__ rz, long long rstride, long long plane, const float* __restrict__ mask,
                         const float* __restrict__ scatter, const float* __restrict__ inv_sigma, float wscat) {
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (p >= plane) return;
    float* r = rz + i * rstride;
    const float v = r[p];
    float z = 0.0f;
    if (mask[p] != 0.0f && finitef(v)) {
        const float s = scatter[p];
        const float e = s > wscat ? s : 1.0f;
        z = v * inv_sigma[i] / e;
    }
    r[p] = z;
}

// Cosmic-ray hits on the interior: z >= hit and the 4-neighbour mean below ratio * z
// (preprocess.cpp:428-440). Counted per frame.
__global__ void hit_kernel(const float* __restrict__ z, long long zstride, int w, int h, float hit, float ratio,
                           unsigned char* __restrict__ flags, int* __restrict__ cr) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const float* zi = z + i * zstride;
    const long long k = static_cast<long long>(y) * w + x;
    unsigned char f = 0;
    if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
        const float v = zi[k];
        if (!(v < hit)) {
            const float n4 = 0.25f * (zi[k - 1] + zi[k + 1] + zi[k - w] + zi[k + w]);
            if (!(n4 >= ratio * v)) {
                f = 1;
                atomicAdd(&cr[i], 1);
            }
        }
    }
    flags[i * (static_cast<long long>(w) * h) + k] = f;
}

// Kill the hits and their 8-neighbours above 2.5 (the kill list is applied after the
// scan, so neighbours are judged on the unkilled z), then clamp to +-clip
// (preprocess.cpp:441-451). Writes to `out` (a separate buffer).
__global__ void kill_clip_kernel(const float* __restrict__ z, long long zstride, int w, int h,
                                 const unsigned char* __restrict__ flags, float clip, float* __restrict__ out,
                                 long long ostride) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const long long k = static_cast<long long>(y) * w + x;
    const long long plane = static_cast<long long>(w) * h;
    const unsigned char* fl = flags + i * plane;
    const float v = z[i * zstride + k];
    bool kill = fl[k] != 0;
    if (!kill && v > 2.5f) {
        for (int dy = -1; dy <= 1 && !kill; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = x + dx, yy = y + dy;
                if ((dx || dy) && xx >= 0 && xx < w && yy >= 0 && yy < h && fl[static_cast<long long>(yy) * w + xx]) {
                    kill = true;
                    break;
                }
            }
    }
    float o = kill ? 0.0f : v;
    o = o < -clip ? -clip : (clip < o ? clip : o);
    out[i * ostride + k] = o;
}

// Separable matched filter with taps truncated at the image edges, multiply then add
// in fp32 in ascending tap order (preprocess.cpp:116-148). taps = float(exp(.)),
// taps_norm = taps * float(1 / sum taps^2), both from the host.
__global__ void mf_rows_kernel(const float* __restrict__ in, long long istride, int w, int R,
                               const float* __restrict__ taps, float* __restrict__ out, long long ostride) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const float* row = in + i * istride + static_cast<long long>(y) * w;
    const int a = max(0, x - R), b = min(w - 1, x + R);
    float acc = 0.0f;
    for (int j = a; j <= b; ++j) acc += row[j] * taps[j - x + R];
    out[i * ostride + static_cast<long long>(y) * w + x] = acc;
}

__global__ void mf_cols_kernel(const float* __restrict__ in, long long istride, int w, int h, int R,
                               const float* __restrict__ taps_norm, float* __restrict__ out, long long ostride) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const float* t = in + i * istride;
    const int a = max(0, y - R), b = min(h - 1, y + R);
    float o = 0.0f;
    for (int j = a; j <= b; ++j) o += taps_norm[j - y + R] * t[static_cast<long long>(j) * w + x];
    out[i * ostride + static_cast<long long>(y) * w + x] = o;
}

// z, cosmic-ray hits, kill + clip and the separable matched filter of one 64 x 32 tile
// of up to `frames` frames in one pass (the chain z_kernel .. mf_cols_kernel above:
// preprocess.cpp:413-451, 116-148), every value computed with the same operations in
// the same order, so the output is the chain's bit for bit. Per tile the mask and
// scatter give each halo pixel its divisor e (scatter where it exceeds weight_scatter,
// else 1; 0 where masked or outside the frame) once for all its frames. Per frame the
// residuals over the tile grown by R + 2 (R the filter half width) become z in LDS; hit
// flags over the tile grown by R + 1 (each needs z at +-1); killed and clipped values
// over the tile grown by R (the kill test reads flags at +-1); the row filter over the
// tile's columns and R more rows each side; the column filter writes the tile. Positions
// outside the frame hold zeros that no in-frame computation reads (the chain's own
// bounds: hits only inside the 1-pixel border, neighbours and taps only inside the
// frame). Hits are counted for the tile's own pixels only, once per pixel. 256 threads
// as 64 columns x 4 rows. LDS: 2 (TX + 2R + 4)(TY + 2R + 4) + (TX + 2R)(TY + 2R) + 2 (2R + 1)
// floats and (TX + 2R + 2)(TY + 2R + 2) bytes; the row-filtered values reuse the z area.
constexpr int kZfTileX = 64, kZfTileY = 32, kZfMaxR = 10;  // R <= 10: at most 62 KB of LDS

__global__ __launch_bounds__(256) void zfilter_kernel(const float* __restrict__ resid, long long rstride, int n,
                                                      int frames, int w, int h, const float* __restrict__ mask,
                                                      const float* __restrict__ scatter,
                                                      const float* __restrict__ inv_sigma, float wscat, float hit,
                                                      float ratio, float clip, const float* __restrict__ taps, int R,
                                                      int* __restrict__ cr, float* __restrict__ out,
                                                      long long ostride) {
    extern __shared__ float zf_lds[];
    constexpr int TX = kZfTileX, TY = kZfTileY;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int i0 = blockIdx.z * frames, i1 = min(n, i0 + frames);
    const int zw = TX + 2 * R + 4, zh = TY + 2 * R + 4;  // e and z: from (x0 - R - 2, y0 - R - 2)
    const int kw = TX + 2 * R, kh = TY + 2 * R;          // killed / clipped: from (x0 - R, y0 - R)
    const int hw = TX + 2 * R + 2, hh = TY + 2 * R + 2;  // hit flags: from (x0 - R - 1, y0 - R - 1)
    const int nt = 2 * R + 1;
    float* const E = zf_lds;
    float* const Z = E + zw * zh;
    float* const K = Z + zw * zh;
    float* const T = K + kw * kh;  // row taps, then the normalised column taps
    unsigned char* const H = reinterpret_cast<unsigned char*>(T + 2 * nt);
    float* const RF = Z;           // row-filtered values, TX x kh (after K is built)
    __shared__ int hits;
    const int tid = threadIdx.x, lx = tid & 63, ly = tid >> 6;
    if (tid == 0) hits = 0;
    for (int e = tid; e < 2 * nt; e += blockDim.x) T[e] = taps[e];
    for (int ey = ly; ey < zh; ey += 4)
        for (int ex = lx; ex < zw; ex += 64) {
            const int gy = y0 - R - 2 + ey, gx = x0 - R - 2 + ex;
            float ee = 0.0f;
            if (gx >= 0 && gx < w && gy >= 0 && gy < h) {
                const long long p = static_cast<long long>(gy) * w + gx;
                if (mask[p] != 0.0f) {
                    const float s = scatter[p];
                    ee = s > wscat ? s : 1.0f;  // >= 1, or +inf
                }
            }
            E[ey * zw + ex] = ee;
        }
    __syncthreads();
    auto zat = [&](int gx, int gy) { return Z[(gy - (y0 - R - 2)) * zw + (gx - (x0 - R - 2))]; };
    auto hat = [&](int gx, int gy) { return H[(gy - (y0 - R - 1)) * hw + (gx - (x0 - R - 1))]; };
    for (int i = i0; i < i1; ++i) {
        const float* r = resid + i * rstride;
        const float inv = inv_sigma[i];
        // z (preprocess.cpp:413-425): r float(1/sigma_i) / e where unmasked and finite, else 0
        for (int ey = ly; ey < zh; ey += 4)
            for (int ex = lx; ex < zw; ex += 64) {
                const float ee = E[ey * zw + ex];
                float z = 0.0f;
                if (ee != 0.0f) {
                    const float v = r[static_cast<long long>(y0 - R - 2 + ey) * w + (x0 - R - 2 + ex)];
                    if (finitef(v)) z = v * inv / ee;
                }
                Z[ey * zw + ex] = z;
            }
        __syncthreads();
        // hits (preprocess.cpp:428-440): interior, z >= hit and the 4-neighbour mean below ratio z
        int mine = 0;
        for (int ey = ly; ey < hh; ey += 4)
            for (int ex = lx; ex < hw; ex += 64) {
                const int gy = y0 - R - 1 + ey, gx = x0 - R - 1 + ex;
                unsigned char f = 0;
                if (gx >= 1 && gx < w - 1 && gy >= 1 && gy < h - 1) {
                    const float v = zat(gx, gy);
                    if (!(v < hit)) {
                        const float n4 =
                            0.25f * (zat(gx - 1, gy) + zat(gx + 1, gy) + zat(gx, gy - 1) + zat(gx, gy + 1));
                        if (!(n4 >= ratio * v)) {
                            f = 1;
                            mine += (gx >= x0 && gx < x0 + TX && gy >= y0 && gy < y0 + TY) ? 1 : 0;
                        }
                    }
                }
                H[ey * hw + ex] = f;
            }
        if (mine) atomicAdd(&hits, mine);
        __syncthreads();
        // kill the hits and their 8-neighbours above 2.5, then clamp to +-clip (:441-451)
        for (int ey = ly; ey < kh; ey += 4)
            for (int ex = lx; ex < kw; ex += 64) {
                const int gy = y0 - R + ey, gx = x0 - R + ex;
                float o = 0.0f;
                if (gx >= 0 && gx < w && gy >= 0 && gy < h) {
                    const float v = zat(gx, gy);
                    bool kill = hat(gx, gy) != 0;
                    if (!kill && v > 2.5f) {
                        for (int dy = -1; dy <= 1 && !kill; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                const int xx = gx + dx, yy = gy + dy;
                                if ((dx || dy) && xx >= 0 && xx < w && yy >= 0 && yy < h && hat(xx, yy)) {
                                    kill = true;
                                    break;
                                }
                            }
                    }
                    o = kill ? 0.0f : v;
                    o = o < -clip ? -clip : (clip < o ? clip : o);
                }
                K[ey * kw + ex] = o;
            }
        __syncthreads();
        if (tid == 0) {
            if (hits) atomicAdd(&cr[i], hits);
            hits = 0;  // the next frame's first hit count comes after two more barriers
        }
        // row filter over the tile's columns and R more rows each side (:116-131): taps
        // truncated at the frame's edges, multiply then add in ascending tap order
        {
            const int gx = x0 + lx;
            const int a = max(0, gx - R), b = min(w - 1, gx + R);
            for (int ey = ly; ey < kh; ey += 4) {
                const int gy = y0 - R + ey;
                float acc = 0.0f;
                if (gx < w && gy >= 0 && gy < h) {
                    const float* krow = K + ey * kw;  // frame column j at krow[j - (x0 - R)]
                    for (int j = a; j <= b; ++j) acc += krow[j - (x0 - R)] * T[j - gx + R];
                }
                RF[ey * TX + lx] = acc;
            }
        }
        __syncthreads();
        // column filter with the normalised taps (:132-148), the tile's pixels
        {
            const int gx = x0 + lx;
            for (int cy = ly; cy < TY; cy += 4) {
                const int gy = y0 + cy;
                if (gx >= w || gy >= h) continue;
                const int a = max(0, gy - R), b = min(h - 1, gy + R);
                float o = 0.0f;
                for (int j = a; j <= b; ++j) o += T[nt + j - gy + R] * RF[(j - (y0 - R)) * TX + lx];
                out[i * ostride + static_cast<long long>(gy) * w + gx] = o;
            }
        }
        __syncthreads();  // the next frame's z overwrites the row-filtered values
    }
}

// g *= float(1 / sg) where sg > 0 (preprocess.cpp:474-477).
__global__ void scale_kernel(float* __restrict__ g, long long gstride, long long plane, const SelectSeg* __restrict__ seg) {
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (p >= plane) return;
    const double sg = seg[i].result;
    if (sg > 0.0) g[i * gstride + p] *= static_cast<float>(1.0 / sg);
}

// scale_kernel's normalisation applied while copying the frames into another layout:
// row y of frame i goes to dst + i dstride + y dpitch, times float(1 / sg) where sg > 0 (the
// same float product as scaling in place, then copying).
__global__ void scale_copy_kernel(const float* __restrict__ src, long long sstride, float* __restrict__ dst,
                                  long long dstride, int dpitch, int w, int y0, const SelectSeg* __restrict__ seg) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = y0 + blockIdx.y, i = blockIdx.z;
    if (x >= w) return;
    const double sg = seg[i].result;
    float v = src[i * sstride + static_cast<long long>(y) * w + x];
    if (sg > 0.0) v *= static_cast<float>(1.0 / sg);
    dst[i * dstride + static_cast<long long>(y) * dpitch + x] = v;
}

// sub = float(v - bg) for finite v, NaN otherwise (preprocess.cpp:233-236)
__global__ void subtract_kernel(const float* __restrict__ raw, long long fstride, long long plane,
                                const SelectSeg* __restrict__ seg, float* __restrict__ sub, long long sstride,
                                unsigned char* __restrict__ bad, int* __restrict__ any_bad) {
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (p >= plane) return;
    const float v = raw[i * fstride + p];
    const float s = finitef(v) ? static_cast<float>(static_cast<double>(v) - seg[i].result) : __builtin_nanf("");
    sub[i * sstride + p] = s;
    const unsigned char b = finitef(s) ? 0 : 1;
    bad[i * plane + p] = b;
    if (b) atomicOr(&any_bad[i], 1);
}

// Grow non-finite areas by `grow` pixels in a square (preprocess.cpp:237-254).
__global__ void grow_kernel(const unsigned char* __restrict__ bad, int w, int h, int grow,
                            const int* __restrict__ any_bad, float* __restrict__ sub, long long sstride) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, i = blockIdx.z;
    if (x >= w || !any_bad[i]) return;
    const long long plane = static_cast<long long>(w) * h;
    const unsigned char* b = bad + i * plane;
    for (int yy = max(0, y - grow); yy <= min(h - 1, y + grow); ++yy)
        for (int xx = max(0, x - grow); xx <= min(w - 1, x + grow); ++xx)
            if (b[static_cast<long long>(yy) * w + xx]) {
                sub[i * sstride + static_cast<long long>(y) * w + x] = __builtin_nanf("");
                return;
            }
}

// ------------------------------------------------------------------ per-frame order statistics
// Samples of segment s (a frame): the regular grid of sample_pixels (common.hpp:85-100),
// finite and (optionally) unmasked values. Keys are ordered float bits; excluded
// samples get 0xffffffff and sort after every finite value. A 4-pass, 8-bit radix
// select finds the order statistics k_lo and k_hi exactly.

__device__ inline bool sample_key(const SelectSeg& s, long long q, const float* __restrict__ mask, int mode,
                                  uint32_t& key) {
    const int gy = static_cast<int>(q / s.nx), gx = static_cast<int>(q % s.nx);
    const int y = s.border + gy * s.step, x = s.border + gx * s.step;
    const long long k = static_cast<long long>(y) * s.w + x;
    const float v = s.img[k];
    if (!finitef(v) || (mask && mask[k] == 0.0f)) {
        key = 0xffffffffu;
        return false;
    }
    // mode 1: absolute deviation from the segment's median, rounded to float (common.hpp:80)
    const float u = mode == 0 ? v : static_cast<float>(__builtin_fabs(static_cast<double>(v) - s.center));
    key = ord_bits(u);
    return true;
}

// Segment i of a uniform frame stack: the sample grid proto on image base + i stride,
// with the selection state cleared (the host computes the grid once for all frames).
__global__ void seg_setup_kernel(SelectSeg* __restrict__ seg, const float* base, long long stride, SelectSeg proto,
                                 int nseg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    SelectSeg s = proto;
    s.img = base + i * stride;
    seg[i] = s;
}

// Counts the valid samples of each segment.
__global__ __launch_bounds__(256) void select_count_kernel(SelectSeg* __restrict__ seg, const float* __restrict__ mask) {
    SelectSeg& s = seg[blockIdx.y];
    const long long G = static_cast<long long>(s.nx) * s.ny;
    int c = 0;
    for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < G;
         q += static_cast<long long>(gridDim.x) * blockDim.x) {
        uint32_t key;
        c += sample_key(s, q, mask, 0, key) ? 1 : 0;
    }
    // wave reduction, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s.count, c);
}

// One 8-bit digit pass of the radix select, counting for both targets of every segment
// (grid x = workgroups per segment, y = segment) the keys whose higher digits equal
// the target's prefix, per digit.
__global__ __launch_bounds__(256) void select_hist_kernel(SelectSeg* __restrict__ seg, const float* __restrict__ mask,
                                                          int mode, int shift, unsigned int* __restrict__ hist) {
    __shared__ unsigned int lh[2][256];
    for (int d = threadIdx.x; d < 512; d += blockDim.x) lh[d >> 8][d & 255] = 0u;
    __syncthreads();
    const SelectSeg& s = seg[blockIdx.y];
    const long long G = static_cast<long long>(s.nx) * s.ny;
    const uint32_t hi_mask = shift >= 24 ? 0u : (0xffffffffu << (shift + 8));
    const uint32_t p0 = s.prefix[0] & hi_mask, p1 = s.prefix[1] & hi_mask;
    for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < G;
         q += static_cast<long long>(gridDim.x) * blockDim.x) {
        uint32_t key;
        sample_key(s, q, mask, mode, key);
        const uint32_t d = (key >> shift) & 255u;
        if ((key & hi_mask) == p0) atomicAdd(&lh[0][d], 1u);
        if ((key & hi_mask) == p1) atomicAdd(&lh[1][d], 1u);
    }
    __syncthreads();
    unsigned int* gh = hist + static_cast<long long>(blockIdx.y) * 512;
    for (int d = threadIdx.x; d < 512; d += blockDim.x) {
        const unsigned int c = lh[d >> 8][d & 255];
        if (c) atomicAdd(&gh[d], c);
    }
}

// The digit of a pass, one workgroup of 256 per segment: inclusive prefix sums of the
// segment's two histograms in LDS, then for each target the digit whose range
// [exclusive, inclusive) holds its rank k (the first digit whose inclusive sum exceeds
// k); when the counts end at or below k, digit 255 and the whole count, as a sequential
// scan would leave them. The histograms are cleared for the next pass.
// mode 0 starts (shift 24) by setting the targets of median_inplace from the count:
// element m/2 and, for even m, element m/2 - 1. After the lowest digit (shift 0) the
// selection finishes: median_inplace (common.hpp:64-74), element m/2 averaged in double
// with element m/2 - 1 for even m, 0 for no samples. mode 0 stores it as the centre (the
// median) and sets the targets again for the deviation pass; mode 1 gives result =
// 1.4826 * median of the deviations (robust_sigma, common.hpp:77-82), 0 when fewer than
// 3 samples, and with `unit_floor` results <= 0 become 1 (sigma_from_samples,
// preprocess.cpp:265-268).
__global__ __launch_bounds__(256) void select_digit_kernel(SelectSeg* __restrict__ seg, int mode, int shift,
                                                           unsigned int* __restrict__ hist, int unit_floor) {
    __shared__ unsigned int lh[2][256];
    __shared__ uint32_t pick[2];
    __shared__ int rem[2];
    SelectSeg& s = seg[blockIdx.x];
    const int m = s.count;
    int kk[2] = {s.k[0], s.k[1]};
    if (mode == 0 && shift == 24) {
        kk[1] = m / 2;
        kk[0] = (m % 2 == 0) ? m / 2 - 1 : m / 2;
        if (kk[0] < 0) kk[0] = 0;
    }
    const int d = threadIdx.x;
    unsigned int* gh = hist + static_cast<long long>(blockIdx.x) * 512;
    const unsigned int h0 = gh[d], h1 = gh[256 + d];
    gh[d] = 0u;
    gh[256 + d] = 0u;
    lh[0][d] = h0;
    lh[1][d] = h1;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned int a0 = d >= off ? lh[0][d - off] : 0u, a1 = d >= off ? lh[1][d - off] : 0u;
        __syncthreads();
        lh[0][d] += a0;
        lh[1][d] += a1;
        __syncthreads();
    }
    for (int t = 0; t < 2; ++t) {
        const unsigned int k = static_cast<unsigned int>(kk[t]);
        const unsigned int incl = lh[t][d], excl = incl - (t ? h1 : h0);
        if (excl <= k && k < incl) {
            pick[t] = static_cast<uint32_t>(d);
            rem[t] = static_cast<int>(k - excl);
        } else if (d == 255 && incl <= k) {
            pick[t] = 255u;
            rem[t] = static_cast<int>(k - incl);
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const uint32_t pre0 = (shift == 24 ? 0u : s.prefix[0]) | (pick[0] << shift);
    const uint32_t pre1 = (shift == 24 ? 0u : s.prefix[1]) | (pick[1] << shift);
    if (shift != 0) {
        s.prefix[0] = pre0;
        s.prefix[1] = pre1;
        s.k[0] = rem[0];
        s.k[1] = rem[1];
        return;
    }
    double med = 0.0;
    if (m > 0) {
        med = static_cast<double>(unord_bits(pre1));
        if (m % 2 == 0) med = 0.5 * (med + static_cast<double>(unord_bits(pre0)));
    }
    s.prefix[0] = s.prefix[1] = 0u;
    if (mode == 0) {
        s.center = med;
        s.result = med;
        s.k[1] = m / 2;
        s.k[0] = (m % 2 == 0) ? m / 2 - 1 : m / 2;
        if (s.k[0] < 0) s.k[0] = 0;
    } else {
        double r = m < 3 ? 0.0 : 1.4826 * med;
        if (unit_floor && !(r > 0.0)) r = 1.0;
        s.result = r;
        s.k[0] = s.k[1] = 0;
    }
}

// 1/sigma as float for the z kernel (preprocess.cpp:416)
__global__ void inv_sigma_kernel(const SelectSeg* __restrict__ seg, int nseg, float* __restrict__ inv,
                                 double* __restrict__ sigma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseg) return;
    sigma[i] = seg[i].result;
    inv[i] = static_cast<float>(1.0 / seg[i].result);
}

}  // namespace ftk

// ------------------------------------------------------------------ host orchestration
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace ftk {

namespace {

constexpr double kFwhmToSigmaD = 1.0 / 2.354820045030949;  // psf.hpp:17

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

template <typename T>
T* dalloc(size_t count, const char* what) {
    T* p = nullptr;
    hcheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(1, count) * sizeof(T)), what);
    return p;
}

// The sample grid of sample_pixels on a w x h image with the given border.
SelectSeg grid_seg(const float* img, int w, int h, int border) {
    SelectSeg s;
    s.img = img;
    s.w = w;
    s.border = border;
    const long long usable = static_cast<long long>(std::max(0, w - 2 * border)) * std::max(0, h - 2 * border);
    s.step = usable == 0 ? 1 : std::max(1, static_cast<int>(std::sqrt(static_cast<double>(usable) / 200000.0)));
    s.nx = std::max(0, (w - 2 * border + s.step - 1) / s.step);
    s.ny = std::max(0, (h - 2 * border + s.step - 1) / s.step);
    if (usable == 0) s.nx = s.ny = 0;
    return s;
}

unsigned blocks_for(long long count, int threads) { return static_cast<unsigned>((count + threads - 1) / threads); }

}  // namespace

DeviceStage::DeviceStage(int n_, int w_, int h_) : n(n_), w(w_), h(h_), plane(static_cast<long long>(w_) * h_) {
    if (n < 1 || w < 3 || h < 3) throw std::runtime_error("build_stage: bad frame stack size");
    // the static-sky kernel keeps each pixel's samples and order in LDS columns (6 n bytes
    // per lane for n > 64) and the scatter map one column (4 n): both within 160 KB
    if (n > kMaxStageFrames)
        throw std::runtime_error("build_stage: at most " + std::to_string(kMaxStageFrames) + " frames on the device path");
    const size_t np = static_cast<size_t>(n) * plane;
    resid_ = dalloc<float>(np, "allocate residuals");
    flags_ = dalloc<unsigned char>(np, "allocate flags");
    sky_ = dalloc<float>(plane, "allocate static sky");
    scatter_ = dalloc<float>(plane, "allocate scatter");
    mask_ = dalloc<float>(plane, "allocate mask");
    seg_ = dalloc<SelectSeg>(n, "allocate selection state");
    hist_ = dalloc<unsigned int>(static_cast<size_t>(n) * 512, "allocate histograms");
    hcheck(hipMemset(hist_, 0, sizeof(unsigned int) * n * 512), "clear histograms");
    // the clear runs on the null stream, which the builds' non-blocking streams do not
    // wait for: it is complete before any of them can read the histograms
    hcheck(hipStreamSynchronize(nullptr), "clear histograms");
    inv_ = dalloc<float>(n, "allocate 1/sigma");
    sigma_ = dalloc<double>(n, "allocate sigma");
    cr_ = dalloc<int>(n, "allocate cosmic-ray counts");
    rank_ = dalloc<int>(n, "allocate ranks");
    any_ = dalloc<int>(n, "allocate flags");
    taps_ = dalloc<float>(2 * 64 + 2, "allocate taps");
}

double DeviceStage::device_bytes() const {
    const double np = static_cast<double>(n) * plane;
    return np * (4.0 + 1.0) + 3.0 * plane * 4.0 + n * (sizeof(SelectSeg) + 512.0 * 4.0 + 4.0 + 8.0 + 4.0 + 4.0 + 4.0) +
           (2.0 * 64 + 2) * 4.0;
}

DeviceStage::~DeviceStage() {
    for (void* p : {static_cast<void*>(resid_), static_cast<void*>(flags_), static_cast<void*>(sky_),
                    static_cast<void*>(scatter_), static_cast<void*>(mask_), static_cast<void*>(seg_),
                    static_cast<void*>(hist_), static_cast<void*>(inv_), static_cast<void*>(sigma_),
                    static_cast<void*>(cr_), static_cast<void*>(rank_), static_cast<void*>(any_),
                    static_cast<void*>(taps_)})
        if (p) (void)hipFree(p);
}

void DeviceStage::select(const float* base, long long stride, int border, const float* mask, bool deviation,
                         int unit_floor, hipStream_t s) {
    const int ns = n;
    const SelectSeg proto = grid_seg(base, w, h, border);
    const int maxg = std::max(1, proto.nx * proto.ny);
    const unsigned gx = std::max(1u, std::min(64u, blocks_for(maxg, 256)));
    const int mode = deviation ? 1 : 0;
    if (!deviation) {
        hipLaunchKernelGGL(seg_setup_kernel, dim3(blocks_for(ns, 64)), dim3(64), 0, s, seg_, base, stride, proto, ns);
        hipLaunchKernelGGL(select_count_kernel, dim3(gx, ns), dim3(256), 0, s, seg_, mask);
    }
    for (int shift = 24; shift >= 0; shift -= 8) {
        hipLaunchKernelGGL(select_hist_kernel, dim3(gx, ns), dim3(256), 0, s, seg_, mask, mode, shift, hist_);
        hipLaunchKernelGGL(select_digit_kernel, dim3(ns), dim3(256), 0, s, seg_, mode, shift, hist_, unit_floor);
    }
    hcheck(hipGetLastError(), "selection kernels");
}

void DeviceStage::background(const float* raw, long long raw_stride, float* sub, long long sub_stride,
                             double* bg_host, hipStream_t s) {
    select(raw, raw_stride, 0, nullptr, false, 0, s);
    hcheck(hipMemsetAsync(any_, 0, sizeof(int) * n, s), "clear flags");
    hipLaunchKernelGGL(subtract_kernel, dim3(blocks_for(plane, 256), n), dim3(256), 0, s, raw, raw_stride, plane, seg_,
                       sub, sub_stride, flags_, any_);
    hipLaunchKernelGGL(grow_kernel, dim3(blocks_for(w, 128), h, n), dim3(128), 0, s, flags_, w, h, 3, any_, sub,
                       sub_stride);
    hcheck(hipGetLastError(), "background kernels");
    if (bg_host) {
        std::vector<SelectSeg> out(static_cast<size_t>(n));
        hcheck(hipMemcpyAsync(out.data(), seg_, sizeof(SelectSeg) * n, hipMemcpyDeviceToHost, s), "download background");
        hcheck(hipStreamSynchronize(s), "background");
        for (int i = 0; i < n; ++i) bg_host[i] = out[static_cast<size_t>(i)].result;
    }
}

void DeviceStage::copy_scaled(const float* src, long long src_stride, float* dst, long long dst_stride, int dst_pitch,
                              hipStream_t s) {
    for (int y0 = 0; y0 < h; y0 += 32768)
        hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(w, 256), std::min(32768, h - y0), n), dim3(256), 0, s, src,
                           src_stride, dst, dst_stride, dst_pitch, w, y0, seg_);
    hcheck(hipGetLastError(), "scale and copy frames");
}

void time_ranks(const double* t_h, int n, int* rank) {
    // stable sort of the frame times (preprocess.cpp:319-331)
    std::vector<int> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t_h[a] < t_h[b]; });
    for (int r = 0; r < n; ++r) rank[order[static_cast<size_t>(r)]] = r;
}

void DeviceStage::build(const float* sub, long long sub_stride, const double* t_h, const StageParams& prm,
                        float* filtered, long long filt_stride, float* static_sky, double* sigma_host,
                        float* mask_out, int* cr_host, hipStream_t s, float* resid_out) {
    std::vector<int> rank(static_cast<size_t>(n));
    time_ranks(t_h, n, rank.data());
    hcheck(hipMemcpyAsync(rank_, rank.data(), sizeof(int) * n, hipMemcpyHostToDevice, s), "upload ranks");
    hcheck(hipStreamSynchronize(s), "upload ranks");  // `rank` is read before it goes
    build_ranked(sub, sub_stride, rank_, prm, filtered, filt_stride, static_sky, sigma_host, mask_out, cr_host, s,
                 resid_out);
}

void DeviceStage::set_taps(const StageParams& prm, hipStream_t s) {
    const double psf_sigma = prm.fwhm_px * kFwhmToSigmaD;
    if (taps_sigma_ == psf_sigma) return;
    const int R = std::max(1, static_cast<int>(std::ceil(3.0 * psf_sigma)));
    if (R > 64) throw std::runtime_error("build_stage: PSF too wide for the device filter (R > 64)");
    std::vector<float> taps(static_cast<size_t>(2 * R + 1)), taps_norm(taps.size());
    double k2 = 0.0;
    for (int i = -R; i <= R; ++i) {
        const double v = std::exp(-0.5 * i * i / (psf_sigma * psf_sigma));
        taps[static_cast<size_t>(i + R)] = static_cast<float>(v);
        k2 += v * v;
    }
    const float norm = static_cast<float>(1.0 / k2);
    for (size_t q = 0; q < taps.size(); ++q) taps_norm[q] = taps[q] * norm;
    std::vector<float> both(taps);
    both.insert(both.end(), taps_norm.begin(), taps_norm.end());
    // earlier builds on this stream may still read the old taps. A copy from pageable
    // memory may return before its DMA has landed, and it runs on the null stream, which
    // non-blocking streams do not wait for: the null stream is synchronised before any
    // build can read the taps
    hcheck(hipStreamSynchronize(s), "build_stage");
    hcheck(hipMemcpy(taps_, both.data(), sizeof(float) * both.size(), hipMemcpyHostToDevice), "upload taps");
    hcheck(hipStreamSynchronize(nullptr), "upload taps");
    taps_sigma_ = psf_sigma;
}

void DeviceStage::build_ranked(const float* sub, long long sub_stride, const int* d_rank, const StageParams& prm,
                               float* filtered, long long filt_stride, float* static_sky, double* sigma_host,
                               float* mask_out, int* cr_host, hipStream_t s, float* resid_out, bool defer_scale) {
    const int exclude = prm.static_exclude >= 0 ? prm.static_exclude : std::max(1, static_cast<int>(std::lround(n / 8.0)));
    const int min_finite = static_cast<int>(std::max<size_t>(5, static_cast<size_t>(0.6 * n)));
    const double psf_sigma = prm.fwhm_px * kFwhmToSigmaD;
    set_taps(prm, s);

    // static sky + leave-out residuals
    float* sky = static_sky ? static_sky : sky_;
    if (n <= 64) {
        // samples in registers, sorted samples and positions in LDS (5 n bytes per lane)
        const size_t lds = static_cast<size_t>(n) * 64 * 5 + 4 * static_cast<size_t>(n);
        auto sky_fn = n <= 16 ? sky_rank_kernel<16> : n <= 32 ? sky_rank_kernel<32> : n <= 48 ? sky_rank_kernel<48>
                                                                                               : sky_rank_kernel<64>;
        hipLaunchKernelGGL(sky_fn, dim3(blocks_for(plane, 64)), dim3(64), lds, s, sub, sub_stride, n, plane, d_rank,
                           exclude, min_finite, sky, resid_, plane);
    } else {
        const size_t lds = static_cast<size_t>(n) * 64 * 6 + 4 * static_cast<size_t>(n);
        hipLaunchKernelGGL(sky_kernel, dim3(blocks_for(plane, 64)), dim3(64), lds, s, sub, sub_stride, n, plane, d_rank,
                           exclude, min_finite, sky, resid_, plane);
    }
    if (resid_out)
        hcheck(hipMemcpyAsync(resid_out, resid_, sizeof(float) * n * plane, hipMemcpyDeviceToDevice, s), "copy residuals");
    // per-frame residual noise: robust sigma of the grid samples, border 2
    select(resid_, plane, 2, nullptr, false, 1, s);
    select(resid_, plane, 2, nullptr, true, 1, s);
    hipLaunchKernelGGL(inv_sigma_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, s, seg_, n, inv_, sigma_);
    // scatter map and mask
    if (n <= 64) {
        auto sc_fn = n <= 16 ? scatter_sort_kernel<16> : n <= 32 ? scatter_sort_kernel<32> : scatter_sort_kernel<64>;
        hipLaunchKernelGGL(sc_fn, dim3(blocks_for(plane, 64)), dim3(64), 0, s, resid_, plane, n, plane, sigma_, scatter_);
    } else {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks_for(plane, 64)), dim3(64), static_cast<size_t>(n) * 64 * 4, s,
                           resid_, plane, n, plane, sigma_, scatter_);
    }
    const int dil = std::max(1, static_cast<int>(std::lround(0.5 * prm.fwhm_px)));
    float* mask = mask_out ? mask_out : mask_;
    hipLaunchKernelGGL(mask_kernel, dim3(blocks_for(w, 128), h), dim3(128), 0, s, scatter_, w, h, dil,
                       prm.mask_scatter, mask);
    // z, cosmic rays, clip
    hcheck(hipMemsetAsync(cr_, 0, sizeof(int) * n, s), "clear counts");
    const float ratio = static_cast<float>(0.5 * std::exp(-0.5 / (psf_sigma * psf_sigma)));
    // matched filter: taps from the host (glibc exp, set_taps), truncated at the edges
    const int R = std::max(1, static_cast<int>(std::ceil(3.0 * psf_sigma)));
    if (R <= kZfMaxR) {
        // z, hits, kill + clip and both filter passes in one tiled pass
        const size_t lds = sizeof(float) * (2 * static_cast<size_t>(kZfTileX + 2 * R + 4) * (kZfTileY + 2 * R + 4) +
                                             static_cast<size_t>(kZfTileX + 2 * R) * (kZfTileY + 2 * R) + 2 * (2 * R + 1)) +
                           static_cast<size_t>(kZfTileX + 2 * R + 2) * (kZfTileY + 2 * R + 2);
        // frames per workgroup: up to 8 (the tile's divisors shared), while the grid keeps
        // at least ~2048 workgroups
        const long long tiles = static_cast<long long>(blocks_for(w, kZfTileX)) * blocks_for(h, kZfTileY);
        const int frames = static_cast<int>(std::max(1LL, std::min<long long>(8, tiles * n / 2048)));
        hipLaunchKernelGGL(zfilter_kernel, dim3(blocks_for(w, kZfTileX), blocks_for(h, kZfTileY), blocks_for(n, frames)),
                           dim3(256), lds, s, resid_, plane, n, frames, w, h, mask, scatter_, inv_,
                           static_cast<float>(prm.weight_scatter), static_cast<float>(prm.cr_sigma), ratio,
                           static_cast<float>(prm.clip_sigma), taps_, R, cr_, filtered, filt_stride);
    } else {
        hipLaunchKernelGGL(z_kernel, dim3(blocks_for(plane, 256), n), dim3(256), 0, s, resid_, plane, plane, mask,
                           scatter_, inv_, static_cast<float>(prm.weight_scatter));
        hipLaunchKernelGGL(hit_kernel, dim3(blocks_for(w, 128), h, n), dim3(128), 0, s, resid_, plane, w, h,
                           static_cast<float>(prm.cr_sigma), ratio, flags_, cr_);
        hipLaunchKernelGGL(kill_clip_kernel, dim3(blocks_for(w, 128), h, n), dim3(128), 0, s, resid_, plane, w, h,
                           flags_, static_cast<float>(prm.clip_sigma), filtered, filt_stride);
        hipLaunchKernelGGL(mf_rows_kernel, dim3(blocks_for(w, 128), h, n), dim3(128), 0, s, filtered, filt_stride, w,
                           R, taps_, resid_, plane);
        hipLaunchKernelGGL(mf_cols_kernel, dim3(blocks_for(w, 128), h, n), dim3(128), 0, s, resid_, plane, w, h, R,
                           taps_ + (2 * R + 1), filtered, filt_stride);
    }
    // normalisation: robust sigma of the masked grid samples, border ceil(3 sigma) + 1
    const int fborder = static_cast<int>(std::ceil(3 * psf_sigma)) + 1;
    select(filtered, filt_stride, fborder, mask, false, 0, s);
    select(filtered, filt_stride, fborder, mask, true, 0, s);
    if (!defer_scale)
        hipLaunchKernelGGL(scale_kernel, dim3(blocks_for(plane, 256), n), dim3(256), 0, s, filtered, filt_stride, plane,
                           seg_);
    hcheck(hipGetLastError(), "build_stage kernels");
    if (sigma_host || cr_host) {
        if (sigma_host) hcheck(hipMemcpyAsync(sigma_host, sigma_, sizeof(double) * n, hipMemcpyDeviceToHost, s), "download sigma");
        if (cr_host) hcheck(hipMemcpyAsync(cr_host, cr_, sizeof(int) * n, hipMemcpyDeviceToHost, s), "download counts");
        hcheck(hipStreamSynchronize(s), "build_stage");
    }
}

}  // namespace ftk
