// gfx950 kernel of the null-peak screen: the artefact tests of the strongest peaks of a
// null run at their grid position and motion (is_artefact_peak, src/search.cpp:229-243:
// the single-frame share of track_snr, edge_flag :167-190, detector_flag :196-225), so
// that artefacts are left out of the pooled null peaks (pool_null_peaks :690-693) as
// the real search rejects them. One wavefront per peak, everything in FP64 with the
// reference's operation order (the build compiles without contraction), so each
// decision is the reference's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "screen.hpp"
#include "stack_launch.hpp"

namespace ftk {
namespace {

constexpr int kWave = 64;

// Frame i's element (x, y) of a pass: fp32, or 16-bit times the frame's power-of-two
// scale (the value the stacking kernel adds).
__device__ inline float frame_at(const ScreenArgs& a, int i, int x, int y) {
    const long long k = static_cast<long long>(i) * a.fstride + static_cast<long long>(y) * a.pitch + x;
    if (a.storage == kStoreF32) return static_cast<const float*>(a.frames)[k];
    const uint16_t q = static_cast<const uint16_t*>(a.frames)[k];
    float v;
    if (a.storage == kStoreBF16) {
        v = __uint_as_float(static_cast<unsigned>(q) << 16);
    } else {
        _Float16 hv;
        __builtin_memcpy(&hv, &q, sizeof q);
        v = static_cast<float>(hv);
    }
    return v * a.scale[i];
}

// bilinear (src/common.hpp:125-138): NaN outside, except exactly on the last row or
// column; the weights and the sum in double, rounded to float once.
__device__ inline float bilinear(const ScreenArgs& a, int i, double x, double y) {
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    if (x0 < 0 || y0 < 0 || x0 + 1 >= a.w || y0 + 1 >= a.h) {
        const int xr = static_cast<int>(lround(x)), yr = static_cast<int>(lround(y));
        if (xr >= 0 && yr >= 0 && xr < a.w && yr < a.h && fabs(x - xr) < 1e-9 && fabs(y - yr) < 1e-9)
            return frame_at(a, i, xr, yr);
        return __builtin_nanf("");
    }
    const double fx = x - x0, fy = y - y0;
    const float r00 = frame_at(a, i, x0, y0), r01 = frame_at(a, i, x0 + 1, y0);
    const float r10 = frame_at(a, i, x0, y0 + 1), r11 = frame_at(a, i, x0 + 1, y0 + 1);
    return static_cast<float>((1 - fy) * ((1 - fx) * r00 + fx * r01) + fy * ((1 - fx) * r10 + fx * r11));
}

__device__ inline double wave_max(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = v < o ? o : v;
    }
    return v;
}

// One workgroup of one wavefront per peak. LDS: n floats (the track's per-frame values).
__global__ void __launch_bounds__(kWave) screen_kernel(ScreenArgs a) {
    extern __shared__ float per_frame[];
    __shared__ double bcast[4];
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.n;
    const double x = a.px[k], y = a.py[k];
    const int vi = a.pv[k];
    const double vx = a.vel[2 * vi], vy = a.vel[2 * vi + 1];

    // track_snr(which = 0) with the per-frame values (search.cpp:139-159)
    for (int i = lane; i < n; i += kWave) {
        const double t = a.t[i];
        per_frame[i] = bilinear(a, i, x + vx * t, y + vy * t);
    }
    __syncthreads();
    if (lane == 0) {
        double sum = 0.0, pos_sum = 0.0, maxc = 0.0;
        int count = 0;
        for (int i = 0; i < n; ++i) {
            const float v = per_frame[i];
            if (!isfinite(v)) continue;
            sum += v;
            ++count;
            const double d = v;
            pos_sum += 0.0 < d ? d : 0.0;  // std::max(0.0, x)
            maxc = maxc < d ? d : maxc;    // std::max(maxc, x)
        }
        const double snr = count >= 3 ? sum / a.sqrtc[count] : 0.0;
        const double frac = pos_sum > 0.0 ? maxc / pos_sum : 1.0;
        bcast[0] = snr;
        bcast[1] = (n >= 8 && frac > a.max_frame_fraction) ? 1.0 : 0.0;
    }
    __syncthreads();
    const double track_snr_all = bcast[0];
    if (bcast[1] != 0.0) {
        if (lane == 0) a.art[k] = 1;
        return;
    }

    // edge_flag: frames whose rounded track position has a blank or outside pixel
    // within r of it (the subtracted frame or the static sky undefined)
    {
        const int r = a.edge_r, side = 2 * r + 1, win = side * side;
        int near_blank = 0;
        for (int i = 0; i < n; ++i) {
            const double t = a.t[i];
            const int xi = static_cast<int>(lround(x + vx * t)), yi = static_cast<int>(lround(y + vy * t));
            const float* sub = a.sub + static_cast<long long>(i) * a.sub_stride;
            bool blank = false;
            for (int e = lane; e < win; e += kWave) {
                const int xx = xi + e % side - r, yy = yi + e / side - r;
                if (xx < 0 || yy < 0 || xx >= a.w || yy >= a.h) {
                    blank = true;
                } else {
                    const long long q = static_cast<long long>(yy) * a.w + xx;
                    if (!isfinite(sub[q]) || !isfinite(a.sky[q])) blank = true;
                }
            }
            near_blank += __any(blank) ? 1 : 0;
        }
        const int lim = n / 10 > 1 ? n / 10 : 1;
        if (near_blank > lim) {
            if (lane == 0) a.art[k] = 1;
            return;
        }
    }

    // detector_flag: the best S/N of a detector-fixed position on the grid of steps
    // around the track's mean detector position, against 0.8 x the track's S/N
    if (a.reg_dx) {
        if (lane == 0) {
            double mx = 0.0, my = 0.0;
            for (int i = 0; i < n; ++i) {
                const double t = a.t[i];
                mx += (x + vx * t) + a.reg_dx[i];
                my += (y + vy * t) + a.reg_dy[i];
            }
            bcast[2] = mx / n;
            bcast[3] = my / n;
        }
        __syncthreads();
        const double mx = bcast[2], my = bcast[3];
        const int ns = a.nsteps, np = ns * ns;
        double best = 0.0;
        for (int p = lane; p < np; p += kWave) {
            const double dxp = mx + a.steps[p % ns], dyp = my + a.steps[p / ns];
            double sum = 0.0;
            int count = 0;
            for (int i = 0; i < n; ++i) {
                const float v = bilinear(a, i, dxp - a.reg_dx[i], dyp - a.reg_dy[i]);
                if (!isfinite(v)) continue;
                sum += v;
                ++count;
            }
            const double f = count >= 3 ? sum / a.sqrtc[count] : 0.0;
            best = best < f ? f : best;
        }
        best = wave_max(best);
        if (lane == 0) a.art[k] = best >= 0.8 * track_snr_all ? 1 : 0;
        return;
    }
    if (lane == 0) a.art[k] = 0;
}

}  // namespace

hipError_t launch_screen(const ScreenArgs& a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * static_cast<size_t>(a.n);
    if (a.n < 1 || lds > 60000) return hipErrorInvalidValue;  // per-frame values of one track in LDS
    hipLaunchKernelGGL(screen_kernel, dim3(static_cast<unsigned>(a.count)), dim3(kWave), lds, s, a);
    return hipGetLastError();
}

}  // namespace ftk
