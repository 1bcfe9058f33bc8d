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
#include "track_device.hpp"

namespace ftk {
namespace {

constexpr int kWave = 64;

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
    // motion j = k nv + iv: velocity iv at inverse distance qs[k] (search.cpp:732-733)
    const int mj = a.pv[k];
    const int vi = a.qs ? mj % a.nv : mj;
    const double q = a.qs ? a.qs[mj / a.nv] : 0.0;
    const double vx = a.vel[2 * vi], vy = a.vel[2 * vi + 1];

    // track_snr(which = 0) with the per-frame values (search.cpp:139-159)
    for (int i = lane; i < n; i += kWave) {
        double tx, ty;
        track_at(f.t, f.par_x, f.par_y, q, i, x, y, vx, vy, tx, ty);
        per_frame[i] = bilinear(f, a.storage, a.w, a.h, i, tx, ty);
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
        edge_rows(a.sub, a.sub_stride, a.sky, a.w, a.h, n, f.t, f.par_x, f.par_y, q, x, y, vx, vy, a.edge_r, lane,
                  kWave, blank);
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
        if (lane == 0) detector_mean(n, f.t, f.par_x, f.par_y, q, a.reg_dx, a.reg_dy, x, y, vx, vy, bcast[2], bcast[3]);
        __syncthreads();
        double best = detector_best(f, a.storage, a.w, a.h, n, a.reg_dx, a.reg_dy, a.steps, a.nsteps, a.sqrtc,
                                    bcast[2], bcast[3], lane, kWave);
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
