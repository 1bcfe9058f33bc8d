// gfx950 kernel of the candidate refinement (refine, src/search.cpp:246-395) for the
// selected peaks of a real pass: the two-level local search over position and motion,
// the parabolic sub-step refinements, the refined track's per-frame samples, S/N and
// halves, the mask count, PSF photometry on the stacked residuals (stack_cutout,
// :826-856) and the field-edge and detector-fixed tests. One workgroup of 256 threads
// per candidate; every S/N is summed by one thread over the frames in order, in FP64 as
// the reference sums it, so each figure and each decision is the reference's.
//
// The local search keeps the first strict maximum of a sequential scan: each thread
// scans its share of the grid in increasing index order with a strict comparison, and
// the reduction takes the larger value, the lower index on a tie, which is the scan's
// result.
#include <hip/hip_runtime.h>

#include <cmath>

#include "refine.hpp"
#include "track_device.hpp"

namespace ftk {
namespace {

constexpr int kRefineThreads = 256;

// std::clamp(v, lo, hi) for lo <= hi
__device__ inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

// The block's maximum of (v, idx) pairs, larger v first, lower idx on a tie (idx -1:
// no grid point beat the start value). Every thread returns the winner.
__device__ inline void block_first_max(double& v, int& idx, double* rv, int* ri) {
    const int tid = threadIdx.x;
    rv[tid] = v;
    ri[tid] = idx;
    __syncthreads();
    for (int s = kRefineThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double v2 = rv[tid + s];
            const int i2 = ri[tid + s];
            const double v1 = rv[tid];
            const int i1 = ri[tid];
            if (v2 > v1 || (v2 == v1 && i2 >= 0 && (i1 < 0 || i2 < i1))) {
                rv[tid] = v2;
                ri[tid] = i2;
            }
        }
        __syncthreads();
    }
    v = rv[0];
    idx = ri[0];
    __syncthreads();
}

__device__ inline double block_max(double v, double* rv) {
    const int tid = threadIdx.x;
    rv[tid] = v;
    __syncthreads();
    for (int s = kRefineThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double o = rv[tid + s];
            rv[tid] = rv[tid] < o ? o : rv[tid];
        }
        __syncthreads();
    }
    const double r = rv[0];
    __syncthreads();
    return r;
}

// Dynamic LDS: n floats (the refined track's samples), n ints (frames blank near the
// track), (2 R + 1)^2 floats (the stacked residual cutout).
__global__ void __launch_bounds__(kRefineThreads) refine_kernel(RefineArgs a) {
    extern __shared__ float lds[];
    const int n = a.n;
    float* const per_frame = lds;
    int* const blank = reinterpret_cast<int*>(lds + n);
    float* const cut = lds + 2 * n;
    __shared__ double rv[kRefineThreads];
    __shared__ int ri[kRefineThreads];
    __shared__ double sb[24];
    __shared__ int counts[2];
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    const ScreenFrames F{a.frames, a.fstride, a.w, 0, nullptr, a.t, a.par_x, a.par_y};
    const ScreenFrames Rs{a.resid, a.fstride, a.w, 0, nullptr, a.t, a.par_x, a.par_y};
    const double* c0 = a.cand + 5 * static_cast<long long>(k);
    double bx = c0[0], by = c0[1], bvx = c0[2], bvy = c0[3], bq = c0[4];
    // S/N of the track at the current inverse distance bq
    auto S = [&](double x, double y, double vx, double vy, int which) {
        return track_snr(F, kStoreF32, a.w, a.h, n, a.sqrtc, x, y, vx, vy, which, bq);
    };
    if (tid == 0) {
        sb[0] = S(bx, by, bvx, bvy, 0);
        counts[0] = counts[1] = 0;
    }
    __syncthreads();
    double best = sb[0];

    // two-level local search around the current best (search.cpp:256-273)
    for (int lv = 0; lv < 2; ++lv) {
        const RefineLevel L = a.lev[lv];
        const double cx = bx, cy = by, cvx = bvx, cvy = bvy;
        const int sv = 2 * L.nv + 1, sp = 2 * L.np + 1;
        const int total = sv * sv * sp * sp;
        double mv = best;
        int mi = -1;
        for (int e = tid; e < total; e += kRefineThreads) {
            // e runs over (a, b, i, j), j fastest, as the reference's nested loops
            const int j = e % sp - L.np;
            const int i = (e / sp) % sp - L.np;
            const int b = (e / (sp * sp)) % sv - L.nv;
            const int av = e / (sp * sp * sv) - L.nv;
            const double vx = cvx + av * L.vstep, vy = cvy + b * L.vstep;
            const double x = cx + i * L.pstep, y = cy + j * L.pstep;
            const double s = S(x, y, vx, vy, 0);
            if (s > mv) {
                mv = s;
                mi = e;
            }
        }
        block_first_max(mv, mi, rv, ri);
        if (mi >= 0) {
            const int j = mi % sp - L.np;
            const int i = (mi / sp) % sp - L.np;
            const int b = (mi / (sp * sp)) % sv - L.nv;
            const int av = mi / (sp * sp * sv) - L.nv;
            best = mv;
            bvx = cvx + av * L.vstep;
            bvy = cvy + b * L.vstep;
            bx = cx + i * L.pstep;
            by = cy + j * L.pstep;
        }
        if (a.dq > 0.0) {
            // near-Earth: the inverse distance in steps of dq L.vstep / dv (dq/4, then
            // dq/8), never negative, in the reference's order (:274-285)
            if (tid == 0) {
                const double step = a.dq * L.vstep / a.dv, cq = bq;
                double bb = best, qb = bq;
                for (int kk = -2; kk <= 2; ++kk) {
                    const double qq = cq + kk * step;
                    if (kk == 0 || qq < 0.0) continue;
                    const double sv2 = track_snr(F, kStoreF32, a.w, a.h, n, a.sqrtc, bx, by, bvx, bvy, 0, qq);
                    if (sv2 > bb) {
                        bb = sv2;
                        qb = qq;
                    }
                }
                sb[15] = bb;
                sb[16] = qb;
            }
            __syncthreads();
            best = sb[15];
            bq = sb[16];
            __syncthreads();
        }
    }

    // sub-pixel position from a parabola through three samples per axis (:287-296)
    {
        const double hh = 0.25;
        if (tid == 0) sb[1] = S(bx - hh, by, bvx, bvy, 0);
        if (tid == 1) sb[2] = S(bx + hh, by, bvx, bvy, 0);
        if (tid == 2) sb[3] = S(bx, by - hh, bvx, bvy, 0);
        if (tid == 3) sb[4] = S(bx, by + hh, bvx, bvy, 0);
        __syncthreads();
        const double sx0 = sb[1], sx2 = sb[2], sy0 = sb[3], sy2 = sb[4];
        const double denx = sx0 - 2 * best + sx2, deny = sy0 - 2 * best + sy2;
        if (denx < 0) bx += clampd(0.5 * hh * (sx0 - sx2) / denx, -hh, hh);
        if (deny < 0) by += clampd(0.5 * hh * (sy0 - sy2) / deny, -hh, hh);
    }
    // sub-step motion from a parabola as well (:299-307)
    {
        const double hh = a.dv / 8;
        if (tid == 0) sb[5] = S(bx, by, bvx - hh, bvy, 0);
        if (tid == 1) sb[6] = S(bx, by, bvx + hh, bvy, 0);
        if (tid == 2) sb[7] = S(bx, by, bvx, bvy - hh, 0);
        if (tid == 3) sb[8] = S(bx, by, bvx, bvy + hh, 0);
        if (tid == 4) sb[9] = S(bx, by, bvx, bvy, 0);
        __syncthreads();
        const double s0 = sb[5], s2 = sb[6], t0 = sb[7], t2 = sb[8], sc = sb[9];
        const double denx = s0 - 2 * sc + s2, deny = t0 - 2 * sc + t2;
        if (denx < 0) bvx += clampd(0.5 * hh * (s0 - s2) / denx, -hh, hh);
        if (deny < 0) bvy += clampd(0.5 * hh * (t0 - t2) / deny, -hh, hh);
    }
    if (a.dq > 0.0) {
        // and the inverse distance (:308-316); std::max(0.0, v) as (0 < v ? v : 0)
        const double hh = a.dq / 8;
        __syncthreads();  // every thread has read sb[5..9]
        if (tid == 0) sb[17] = S(bx, by, bvx, bvy, 0);
        if (tid == 1) {
            const double lo = bq - hh;
            sb[18] = track_snr(F, kStoreF32, a.w, a.h, n, a.sqrtc, bx, by, bvx, bvy, 0, 0.0 < lo ? lo : 0.0);
        }
        if (tid == 2) sb[19] = track_snr(F, kStoreF32, a.w, a.h, n, a.sqrtc, bx, by, bvx, bvy, 0, bq + hh);
        __syncthreads();
        const double sc = sb[17], s0 = sb[18], s2 = sb[19];
        const double den = s0 - 2 * sc + s2;
        if (den < 0 && bq - hh >= 0.0) {
            const double nq = bq + clampd(0.5 * hh * (s0 - s2) / den, -hh, hh);
            bq = 0.0 < nq ? nq : 0.0;
        }
    }

    // the refined track's samples, and the frames where it is inside and masked (:318-368)
    int inside = 0, masked = 0;
    for (int i = tid; i < n; i += kRefineThreads) {
        double px, py;
        track_at(a.t, a.par_x, a.par_y, bq, i, bx, by, bvx, bvy, px, py);
        const float v = bilinear(F, kStoreF32, a.w, a.h, i, px, py);
        per_frame[i] = v;
        blank[i] = 0;
        if (isfinite(v)) {
            ++inside;
            const int xi = static_cast<int>(lround(px)), yi = static_cast<int>(lround(py));
            if (xi >= 0 && yi >= 0 && xi < a.w && yi < a.h && a.mask[static_cast<long long>(yi) * a.w + xi] == 0.0f)
                ++masked;
        }
    }
    if (inside) atomicAdd(&counts[0], inside);
    if (masked) atomicAdd(&counts[1], masked);
    __syncthreads();

    // S/N and halves (:318, 346-347); field-edge rows; the stacked residual cutout
    // (:365-377); the detector test's mean position
    if (tid == 0) {
        double sum = 0.0;
        int c = 0;
        for (int i = 0; i < n; ++i) {
            const float v = per_frame[i];
            if (!isfinite(v)) continue;
            sum += v;
            ++c;
        }
        sb[10] = c >= 3 ? sum / a.sqrtc[c] : 0.0;
    }
    if (tid == 1) sb[11] = S(bx, by, bvx, bvy, 1);
    if (tid == 2) sb[12] = S(bx, by, bvx, bvy, 2);
    if (tid == 3 && a.reg_dx)
        detector_mean(n, a.t, a.par_x, a.par_y, bq, a.reg_dx, a.reg_dy, bx, by, bvx, bvy, sb[13], sb[14]);
    edge_rows(a.sub, a.fstride, a.sky, a.w, a.h, n, a.t, a.par_x, a.par_y, bq, bx, by, bvx, bvy, a.edge_r, tid,
              kRefineThreads, blank);
    const int size = 2 * a.R + 1;
    for (int p = tid; p < size * size; p += kRefineThreads) {
        const int u = p % size, v = p / size;
        float acc = 0.0f;
        int c = 0;
        for (int i = 0; i < n; ++i) {
            double cx, cy;
            track_at(a.t, a.par_x, a.par_y, bq, i, bx, by, bvx, bvy, cx, cy);
            const float s = bilinear(Rs, kStoreF32, a.w, a.h, i, cx + u - a.R, cy + v - a.R);
            if (!isfinite(s)) continue;
            acc += s;
            ++c;
        }
        cut[p] = c == 0 ? __builtin_nanf("") : acc / static_cast<float>(c);
    }
    __syncthreads();
    const double snr = sb[10];

    // detector-fixed scan, then the sums of the thread-0 figures
    double det = 0.0;
    if (a.reg_dx) {
        const double bestf = detector_best(F, kStoreF32, a.w, a.h, n, a.reg_dx, a.reg_dy, a.steps, a.nsteps, a.sqrtc,
                                           sb[13], sb[14], tid, kRefineThreads);
        det = block_max(bestf, rv) >= 0.8 * snr ? 1.0 : 0.0;
    }
    if (tid == 0) {
        int near_blank = 0;
        for (int i = 0; i < n; ++i) near_blank += blank[i];
        const int lim = n / 10 > 1 ? n / 10 : 1;
        double num = 0.0, den = 0.0;
        for (int p = 0; p < size * size; ++p) {
            const float v = cut[p];
            if (!isfinite(v)) continue;
            const double kw = a.kern[p];
            num += v * kw;
            den += kw * kw;
        }
        double* o = a.out + static_cast<long long>(k) * kRefOut;
        o[kRefX] = bx;
        o[kRefY] = by;
        o[kRefVx] = bvx;
        o[kRefVy] = bvy;
        o[kRefSnr] = snr;
        o[kRefHalf1] = sb[11];
        o[kRefHalf2] = sb[12];
        o[kRefNum] = num;
        o[kRefDen] = den;
        o[kRefInside] = counts[0];
        o[kRefMasked] = counts[1];
        o[kRefEdge] = n > 0 && near_blank > lim ? 1.0 : 0.0;
        o[kRefDetector] = det;
        o[kRefQ] = bq;
    }
    for (int i = tid; i < n; i += kRefineThreads) {
        const float v = per_frame[i];
        a.contrib[static_cast<long long>(k) * n + i] = isfinite(v) ? v : __builtin_nanf("");
    }
}

}  // namespace

hipError_t launch_refine(const RefineArgs& a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    const long long size = 2LL * a.R + 1;
    const size_t lds = sizeof(float) * (2 * static_cast<size_t>(a.n) + static_cast<size_t>(size * size));
    if (a.n < 1 || a.R < 0 || lds > 60000) return hipErrorInvalidValue;
    hipLaunchKernelGGL(refine_kernel, dim3(static_cast<unsigned>(a.count)), dim3(kRefineThreads), lds, s, a);
    return hipGetLastError();
}

}  // namespace ftk
