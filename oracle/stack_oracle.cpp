// CPU oracle: the stacking contract (TEST INFRASTRUCTURE ONLY; see oracle.h).
//
// Restates include/fainttrail/backend.hpp:10-40 (the contract), src/backend.cpp:128-172
// (naive), src/backend_cpu.cpp:46-204 (cache-blocked CPU backend) and
// src/backend.cpp:281-314 (check_backend). Written from the contract text; the
// blocked form uses its own coverage bookkeeping (a per-row difference array over
// the frames' column intervals), which yields the same counts as the reference's
// sorted edge walk.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

namespace {

inline int nthreads(int requested) {
#ifdef _OPENMP
    return requested > 0 ? requested : omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

// 1/sqrt(c) as float from the double quotient (backend_cpu.cpp:103-104).
std::vector<float> inv_sqrt_table(int n) {
    std::vector<float> t(static_cast<size_t>(n) + 1, 0.0f);
    for (int c = 1; c <= n; ++c) t[static_cast<size_t>(c)] = static_cast<float>(1.0 / std::sqrt(static_cast<double>(c)));
    return t;
}

// S_v at one pixel per the contract; false when fewer than min_count frames cover it
// (backend.cpp:131-146).
bool snr_at(int n, int w, int h, const float* frames, const int32_t* dx, const int32_t* dy, int j, int x, int y,
            int min_count, float& out) {
    float sum = 0.0f;
    int c = 0;
    const size_t plane = static_cast<size_t>(w) * h;
    for (int i = 0; i < n; ++i) {
        const int sx = x + dx[static_cast<size_t>(j) * n + i];
        const int sy = y + dy[static_cast<size_t>(j) * n + i];
        if (sx < 0 || sy < 0 || sx >= w || sy >= h) continue;
        sum += frames[static_cast<size_t>(i) * plane + static_cast<size_t>(sy) * w + sx];
        ++c;
    }
    if (c < min_count) return false;
    out = sum * static_cast<float>(1.0 / std::sqrt(static_cast<double>(c)));
    return true;
}

}  // namespace

extern "C" int ora_stack_naive(int n, int w, int h, const float* frames, int nv, const int32_t* dx,
                               const int32_t* dy, int min_count, float* max_snr, int32_t* best_v) {
    if (n < 1 || w < 1 || h < 1 || nv < 1 || min_count < 1) return 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float best = -std::numeric_limits<float>::infinity();
            int arg = -1;
            for (int j = 0; j < nv; ++j) {
                float s;
                if (!snr_at(n, w, h, frames, dx, dy, j, x, y, min_count, s)) continue;
                if (s > best) {  // strict: equal values keep the lower motion index
                    best = s;
                    arg = j;
                }
            }
            const size_t o = static_cast<size_t>(y) * w + x;
            max_snr[o] = arg >= 0 ? best : 0.0f;
            best_v[o] = arg;
        }
    return 0;
}

namespace {

struct Rect {
    int dx, dy;
    int x0, x1, y0, y1;  // output pixels covered by the frame, [x0,x1) x [y0,y1); empty when x0 == x1
};

// Automatic block sizes (backend_cpu.cpp:46-65): 8-row bands halved until there are
// at least 4 per thread; motions per block so that the accumulators take ~640 KiB.
void auto_blocking(int w, int h, int nv, int nt, int& band, int& vblock) {
    if (band <= 0) {
        band = 8;
        while (band > 1 && (h + band - 1) / band < 4 * nt) band = std::max(1, band / 2);
    }
    band = std::clamp(band, 1, std::max(1, h));
    if (vblock <= 0) {
        const double per_motion = 4.0 * band * std::max(1, w);
        vblock = static_cast<int>(std::clamp(640.0 * 1024 / per_motion, 1.0, 16.0));
    }
    vblock = std::clamp(vblock, 1, std::max(1, nv));
}

// frames holds rows [fr0, fr0 + stride / w) of every frame, frame i at frames + i * stride
int stack_blocked(int n, int w, int h, const float* frames, size_t stride, int fr0, int nv, const int32_t* dx,
                  const int32_t* dy, int min_count, int threads, int band, int vblock, int ya_all, int yb_all,
                  float* max_snr, int32_t* best_v) {
    if (n < 1 || w < 1 || h < 1 || nv < 1 || min_count < 1) return 1;
    const int nt = nthreads(threads);
    auto_blocking(w, yb_all - ya_all, nv, nt, band, vblock);
    std::vector<Rect> rect(static_cast<size_t>(nv) * n);
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < n; ++i) {
            Rect& r = rect[static_cast<size_t>(j) * n + i];
            r.dx = dx[static_cast<size_t>(j) * n + i];
            r.dy = dy[static_cast<size_t>(j) * n + i];
            r.x0 = std::max(0, -r.dx);
            r.x1 = std::min(w, w - r.dx);
            r.y0 = std::max(0, -r.dy);
            r.y1 = std::min(h, h - r.dy);
            if (r.x0 >= r.x1 || r.y0 >= r.y1) r.x0 = r.x1 = r.y0 = r.y1 = 0;
        }
    const std::vector<float> inv = inv_sqrt_table(n);
    const float neg = -std::numeric_limits<float>::infinity();
    const int rows_out = yb_all - ya_all;
    for (size_t q = 0; q < static_cast<size_t>(rows_out) * w; ++q) {
        max_snr[q] = neg;
        best_v[q] = -1;
    }
    const int nbands = (rows_out + band - 1) / band;
#pragma omp parallel num_threads(nt)
    {
        std::vector<float> acc(static_cast<size_t>(vblock) * band * w);
        std::vector<int> diff(static_cast<size_t>(w) + 1);
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < nbands; ++b) {
            const int ya = ya_all + b * band, yb = std::min(yb_all, ya + band), bh = yb - ya;
            for (int v0 = 0; v0 < nv; v0 += vblock) {
                const int vn = std::min(vblock, nv - v0);
                std::fill(acc.begin(), acc.begin() + static_cast<ptrdiff_t>(vn) * bh * w, 0.0f);
                // frames in ascending order; each motion's rows get the shifted frame rows
                for (int i = 0; i < n; ++i) {
                    const float* g = frames + static_cast<size_t>(i) * stride - static_cast<ptrdiff_t>(fr0) * w;
                    for (int k = 0; k < vn; ++k) {
                        const Rect& r = rect[static_cast<size_t>(v0 + k) * n + i];
                        const int y0 = std::max(ya, r.y0), y1 = std::min(yb, r.y1);
                        for (int y = y0; y < y1; ++y) {
                            float* o = &acc[(static_cast<size_t>(k) * bh + (y - ya)) * w];
                            const float* s = g + static_cast<ptrdiff_t>(y + r.dy) * w + r.dx;
                            for (int x = r.x0; x < r.x1; ++x) o[x] += s[x];
                        }
                    }
                }
                // coverage of each row: +1/-1 at the column interval ends of the frames
                // that reach the row, then a prefix sum; motions in ascending order with a
                // strict '>' keep the lowest index on ties
                for (int k = 0; k < vn; ++k) {
                    const int j = v0 + k;
                    const Rect* R = &rect[static_cast<size_t>(j) * n];
                    for (int y = ya; y < yb; ++y) {
                        std::fill(diff.begin(), diff.end(), 0);
                        int frames_on_row = 0;
                        for (int i = 0; i < n; ++i)
                            if (y >= R[i].y0 && y < R[i].y1) {
                                ++diff[static_cast<size_t>(R[i].x0)];
                                --diff[static_cast<size_t>(R[i].x1)];
                                ++frames_on_row;
                            }
                        if (frames_on_row < min_count) continue;
                        const float* arow = &acc[(static_cast<size_t>(k) * bh + (y - ya)) * w];
                        float* mrow = &max_snr[static_cast<size_t>(y - ya_all) * w];
                        int32_t* brow = &best_v[static_cast<size_t>(y - ya_all) * w];
                        int c = 0;
                        for (int x = 0; x < w; ++x) {
                            c += diff[static_cast<size_t>(x)];
                            if (c < min_count) continue;
                            const float snr = arow[x] * inv[static_cast<size_t>(c)];
                            if (snr > mrow[x]) {
                                mrow[x] = snr;
                                brow[x] = j;
                            }
                        }
                    }
                }
            }
        }
    }
    // undefined pixels (-inf) and any non-finite maximum become 0 (backend_cpu.cpp:191-192)
    for (size_t q = 0; q < static_cast<size_t>(rows_out) * w; ++q)
        if (!std::isfinite(max_snr[q])) max_snr[q] = 0.0f;
    return 0;
}

}  // namespace

extern "C" int ora_stack_blocked(int n, int w, int h, const float* frames, int nv, const int32_t* dx,
                                 const int32_t* dy, int min_count, int threads, int band_rows, int vel_block,
                                 float* max_snr, int32_t* best_v) {
    return stack_blocked(n, w, h, frames, static_cast<size_t>(w) * h, 0, nv, dx, dy, min_count, threads, band_rows,
                         vel_block, 0, h, max_snr, best_v);
}

extern "C" int ora_stack_blocked_rows(int n, int w, int h, const float* frames, int fr0, int fr1, int nv,
                                      const int32_t* dx, const int32_t* dy, int min_count, int threads, int y0, int y1,
                                      float* max_snr, int32_t* best_v) {
    if (y0 < 0 || y1 > h || y0 >= y1 || fr0 < 0 || fr1 > h || fr0 >= fr1) return 1;
    // every frame row an output row in [y0, y1) can reach must be present
    for (size_t q = 0; q < static_cast<size_t>(nv) * n; ++q) {
        const int lo = y0 + dy[q], hi = y1 - 1 + dy[q];
        if (hi >= 0 && lo < h && (std::max(lo, 0) < fr0 || std::min(hi, h - 1) >= fr1)) return 2;
    }
    return stack_blocked(n, w, h, frames, static_cast<size_t>(fr1 - fr0) * w, fr0, nv, dx, dy, min_count, threads, 0,
                         0, y0, y1, max_snr, best_v);
}

extern "C" void ora_linear_shifts(int n, const double* t_h, int nv, const double* vxy, int32_t* dx, int32_t* dy) {
    for (int j = 0; j < nv; ++j)
        for (int i = 0; i < n; ++i) {
            dx[static_cast<size_t>(j) * n + i] = static_cast<int32_t>(std::lround(vxy[2 * j] * t_h[i]));
            dy[static_cast<size_t>(j) * n + i] = static_cast<int32_t>(std::lround(vxy[2 * j + 1] * t_h[i]));
        }
}

extern "C" void ora_check_backend(int n, int w, int h, const float* frames, int nv, const int32_t* dx,
                                  const int32_t* dy, int min_count, const float* ref_max, const int32_t* ref_best,
                                  const float* cand_max, const int32_t* cand_best, double* out) {
    double max_abs = 0.0;
    size_t bad = 0, mism = 0, unexpl = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t k = static_cast<size_t>(y) * w + x;
            const double a = ref_max[k], b = cand_max[k];
            const double tol = 1e-4 * std::max(1.0, std::fabs(a));
            const double d = std::fabs(a - b);
            if (!(d <= tol)) ++bad;
            if (std::isfinite(d)) max_abs = std::max(max_abs, d);
            const int va = ref_best[k], vb = cand_best[k];
            if (va == vb) continue;
            ++mism;
            float sa = 0, sb = 0;
            const bool da = va >= 0 && va < nv && snr_at(n, w, h, frames, dx, dy, va, x, y, min_count, sa);
            const bool db = vb >= 0 && vb < nv && snr_at(n, w, h, frames, dx, dy, vb, x, y, min_count, sb);
            if (!(da && db && std::fabs(sa - sb) <= tol)) ++unexpl;
        }
    out[0] = (bad == 0 && unexpl == 0) ? 1.0 : 0.0;
    out[1] = max_abs;
    out[2] = static_cast<double>(bad);
    out[3] = static_cast<double>(mism);
    out[4] = static_cast<double>(unexpl);
}
