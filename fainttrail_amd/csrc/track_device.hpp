// Device restatements of the reference's per-track evaluations, shared by the null-peak
// screen and the candidate refinement: bilinear sampling (src/common.hpp:125-138), the
// track S/N (track_snr, src/search.cpp:142-162), the field-edge window test (edge_flag,
// :167-190) and the detector-fixed scan (detector_flag, :196-225). Everything in FP64
// with the reference's operation order; the build compiles without contraction, so a
// result is the reference's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "screen.hpp"
#include "stack_launch.hpp"

namespace ftk {

// Frame i's element (x, y): fp32, or 16-bit times the frame's power-of-two scale (the
// value the stacking kernel adds).
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

// track_pos (src/internal.hpp:107-116): frame i's position of the track through (x, y)
// at t = 0 with motion (vx, vy) and inverse distance q (a curved track when q != 0:
// plus q times the frame's parallax displacement).
__device__ inline void track_at(const double* t_h, const double* par_x, const double* par_y, double q, int i,
                                double x, double y, double vx, double vy, double& ox, double& oy) {
    const double t = t_h[i];
    ox = x + vx * t;
    oy = y + vy * t;
    if (q != 0.0) {
        ox += q * par_x[i];
        oy += q * par_y[i];
    }
}

// track_snr without the per-frame values, by one thread: the frames in order (which: 0
// all, 1 those with t < 0, 2 those with t >= 0), the finite samples summed in double;
// sqrtc[c] = sqrt(c).
__device__ inline double track_snr(const ScreenFrames& f, int storage, int w, int h, int n, const double* sqrtc,
                                   double x, double y, double vx, double vy, int which, double q) {
    double sum = 0.0;
    int count = 0;
    for (int i = 0; i < n; ++i) {
        const double t = f.t[i];
        if (which == 1 && t >= 0.0) continue;
        if (which == 2 && t < 0.0) continue;
        double px, py;
        track_at(f.t, f.par_x, f.par_y, q, i, x, y, vx, vy, px, py);
        const float v = bilinear(f, storage, w, h, i, px, py);
        if (!isfinite(v)) continue;
        sum += v;
        ++count;
    }
    return count >= 3 ? sum / sqrtc[count] : 0.0;
}

// edge_flag's window test, spread over nl threads: thread lane takes the window rows
// lane, lane + nl, ... (one row of one frame each, so its loads are independent) and
// sets blank[i] for a frame whose rounded track position has an outside pixel, or one
// where the subtracted frame or the static sky is undefined, within r of it. blank
// (n ints) must be zero on entry; the caller synchronises before reading it.
__device__ inline void edge_rows(const float* sub, long long sub_stride, const float* sky, int w, int h, int n,
                                 const double* t_h, const double* par_x, const double* par_y, double q, double x,
                                 double y, double vx, double vy, int r, int lane, int nl, int* blank) {
    const int side = 2 * r + 1;
    for (int e = lane; e < n * side; e += nl) {
        const int i = e / side;
        double px, py;
        track_at(t_h, par_x, par_y, q, i, x, y, vx, vy, px, py);
        const int xi = static_cast<int>(lround(px));
        const int yy = static_cast<int>(lround(py)) + (e - i * side) - r;
        // a row that leaves the frame holds an outside pixel
        int any = yy < 0 || yy >= h || xi - r < 0 || xi + r >= w;
        if (!any) {
            const long long o = static_cast<long long>(yy) * w + (xi - r);
            const float* sr = sub + static_cast<long long>(i) * sub_stride + o;
            const float* kr = sky + o;
            // no short cut: the loads stay independent
#pragma unroll 8
            for (int d = 0; d < side; ++d) any |= static_cast<int>(!isfinite(sr[d])) | static_cast<int>(!isfinite(kr[d]));
        }
        if (any) blank[i] = 1;
    }
}

// The track's mean detector position (detector_flag's mx, my), by one thread.
__device__ inline void detector_mean(int n, const double* t_h, const double* par_x, const double* par_y, double q,
                                     const double* reg_dx, const double* reg_dy, double x, double y, double vx,
                                     double vy, double& mx, double& my) {
    mx = 0.0;
    my = 0.0;
    for (int i = 0; i < n; ++i) {
        double px, py;
        track_at(t_h, par_x, par_y, q, i, x, y, vx, vy, px, py);
        mx += px + reg_dx[i];
        my += py + reg_dy[i];
    }
    mx /= n;
    my /= n;
}

// detector_flag's scan, spread over nl threads: the best S/N of a detector-fixed
// position on the grid mx + steps[a], my + steps[b] among the positions lane, lane + nl,
// ...; the caller takes the maximum over threads and compares it with 0.8 x the track's
// S/N.
__device__ inline double detector_best(const ScreenFrames& f, int storage, int w, int h, int n, const double* reg_dx,
                                       const double* reg_dy, const double* steps, int ns, const double* sqrtc,
                                       double mx, double my, int lane, int nl) {
    double best = 0.0;
    for (int p = lane; p < ns * ns; p += nl) {
        const double dxp = mx + steps[p % ns], dyp = my + steps[p / ns];
        double sum = 0.0;
        int count = 0;
        for (int i = 0; i < n; ++i) {
            const float v = bilinear(f, storage, w, h, i, dxp - reg_dx[i], dyp - reg_dy[i]);
            if (!isfinite(v)) continue;
            sum += v;
            ++count;
        }
        const double s = count >= 3 ? sum / sqrtc[count] : 0.0;
        best = best < s ? s : best;
    }
    return best;
}

}  // namespace ftk
