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
__device__ inline float frame_at(const ScreenFrames& f, int storage, int i, int x, int y) {
    const long long k = static_cast<long long>(i) * f.fstride + static_cast<long long>(y) * f.pitch + x;
    if (storage == kStoreF32) return static_cast<const float*>(f.frames)[k];
    const uint16_t q = static_cast<const uint16_t*>(f.frames)[k];
    float v;
    if (storage == kStoreBF16) {
        v = __uint_as_float(static_cast<unsigned>(q) << 16);
    } else {
        _Float16 hv;
        __builtin_memcpy(&hv, &q, sizeof q);
        v = static_cast<float>(hv);
    }
    return v * f.scale[i];
}

// bilinear (src/common.hpp:125-138): NaN outside, except exactly on the last row or
// column; the weights and the sum in double, rounded to float once.
__device__ inline float bilinear(const ScreenFrames& f, int storage, int w, int h, int i, double x, double y) {
    const int x0 = static_cast<int>(floor(x)), y0 = static_cast<int>(floor(y));
    if (x0 < 0 || y0 < 0 || x0 + 1 >= w || y0 + 1 >= h) {
        const int xr = static_cast<int>(lround(x)), yr = static_cast<int>(lround(y));
        if (xr >= 0 && yr >= 0 && xr < w && yr < h && fabs(x - xr) < 1e-9 && fabs(y - yr) < 1e-9)
            return frame_at(f, storage, i, xr, yr);
        return __builtin_nanf("");
    }
    const double fx = x - x0, fy = y - y0;
    const float r00 = frame_at(f, storage, i, x0, y0), r01 = frame_at(f, storage, i, x0 + 1, y0);
    const float r10 = frame_at(f, storage, i, x0, y0 + 1), r11 = frame_at(f, storage, i, x0 + 1, y0 + 1);
    return static_cast<float>((1 - fy) * ((1 - fx) * r00 + fx * r01) + fy * ((1 - fx) * r10 + fx * r11));
}

__device__ inline double wave_max(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, kWave);
        v = v < o ? o : v;
    }
    return v;
}

// One workgroup of one wavefront per peak. LDS: n floats (the track's per-frame values)
// and n ints (the frames found blank near the track).
__global__ void __launch_bounds__(kWave) screen_kernel(ScreenArgs a) {
    extern __shared__ float per_frame[];
    int* const blank = reinterpret_cast<int*>(per_frame + a.n);
    __shared__ double bcast[4];
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    const int n = a.n;
    const ScreenFrames f = a.passes[a.pslot[k]];
    const double x = a.px[k], y = a.py[k];
    const int vi = a.pv[k];
    const double vx = a.vel[2 * vi], vy = a.vel[2 * vi + 1];

    // track_snr(which = 0) with the per-frame values (search.cpp:139-159)
    for (int i = lane; i < n; i += kWave) {
        const double t = f.t[i];
        per_frame[i] = bilinear(f, a.storage, a.w, a.h, i, x + vx * t, y + vy * t);
        blank[i] = 0;
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
    // within r of it (the subtracted frame or the static sky undefined); a lane takes
    // one window row of one frame, so its loads are independent of each other
    {
        const int r = a.edge_r, side = 2 * r + 1;
        for (int e = lane; e < n * side; e += kWave) {
            const int i = e / side;
            const double t = f.t[i];
            const int xi = static_cast<int>(lround(x + vx * t));
            const int yy = static_cast<int>(lround(y + vy * t)) + (e - i * side) - r;
            // a row that leaves the frame holds an outside pixel
            int any = yy < 0 || yy >= a.h || xi - r < 0 || xi + r >= a.w;
            if (!any) {
                const long long o = static_cast<long long>(yy) * a.w + (xi - r);
                const float* sr = a.sub + static_cast<long long>(i) * a.sub_stride + o;
                const float* kr = a.sky + o;
                // no short cut: the loads stay independent
#pragma unroll 8
                for (int d = 0; d < side; ++d) any |= static_cast<int>(!isfinite(sr[d])) | static_cast<int>(!isfinite(kr[d]));
            }
            if (any) blank[i] = 1;
        }
        __syncthreads();
        int near_blank = 0;
        for (int i = lane; i < n; i += kWave) near_blank += blank[i];
        for (int off = kWave / 2; off > 0; off >>= 1) near_blank += __shfl_xor(near_blank, off, kWave);
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
                const double t = f.t[i];
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
                const float v = bilinear(f, a.storage, a.w, a.h, i, dxp - a.reg_dx[i], dyp - a.reg_dy[i]);
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
    const size_t lds = (sizeof(float) + sizeof(int)) * static_cast<size_t>(a.n);
    if (a.n < 1 || lds > 60000) return hipErrorInvalidValue;  // per-frame values of one track in LDS
    hipLaunchKernelGGL(screen_kernel, dim3(static_cast<unsigned>(a.count)), dim3(kWave), lds, s, a);
    return hipGetLastError();
}

}  // namespace ftk
