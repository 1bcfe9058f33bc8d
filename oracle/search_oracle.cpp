// CPU oracle: search-side helpers around the stacking pass (TEST INFRASTRUCTURE ONLY;
// see oracle.h). Restates src/search.cpp:39-58 (velocity_grid), :105-135 (find_peaks),
// :471-612 (tail fits, threshold), :559-571 (null_time_orders), src/random.hpp:14-85
// (xoshiro256** with splitmix64 seeding; Blackman & Vigna's published generator) and
// the reference tests' xorshift64 + Box-Muller generator (tests/test_main.cpp:28-36).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <numeric>
#include <vector>

#include "oracle.h"

namespace {

inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

}  // namespace

extern "C" void ora_test_normals(uint64_t seed, size_t count, float* out) {
    uint64_t s = seed;
    auto uniform = [&s]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (static_cast<double>(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    for (size_t k = 0; k < count; ++k) {
        const double u1 = uniform();
        const double u2 = uniform();
        out[k] = static_cast<float>(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
}

extern "C" void ora_rng_seed(ora_rng* r, uint64_t seed) {
    uint64_t z = seed;
    for (auto& st : r->s) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t q = z;
        q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ULL;
        q = (q ^ (q >> 27)) * 0x94d049bb133111ebULL;
        st = q ^ (q >> 31);
    }
    r->have_spare = 0;
    r->spare = 0.0;
}

extern "C" uint64_t ora_rng_next(ora_rng* r) {
    uint64_t* s = r->s;
    const uint64_t out = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

extern "C" double ora_rng_uniform(ora_rng* r) {
    return static_cast<double>(ora_rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

// Marsaglia's polar method with a cached second deviate (random.hpp:39-52).
extern "C" double ora_rng_normal(ora_rng* r) {
    if (r->have_spare) {
        r->have_spare = 0;
        return r->spare;
    }
    double u, v, s;
    do {
        u = ora_rng_uniform(r) * 2.0 - 1.0;
        v = ora_rng_uniform(r) * 2.0 - 1.0;
        s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    const double m = std::sqrt(-2.0 * std::log(s) / s);
    r->spare = v * m;
    r->have_spare = 1;
    return u * m;
}

extern "C" int ora_velocity_grid(double span_h, double vmin, double vmax, double tol, double* vxy, int cap,
                                 double* dv_out) {
    if (!(span_h > 0.0) || vmax <= 0.0 || vmax < vmin) return -1;
    const double dv = 4.0 * std::max(0.05, tol) / span_h;
    if (dv_out) *dv_out = dv;
    const int K = static_cast<int>(std::ceil(vmax / dv));
    const double lo = std::max(0.0, vmin - 0.5 * dv), hi = vmax + 0.5 * dv;
    long count = 0;
    for (int j = -K; j <= K; ++j)
        for (int i = -K; i <= K; ++i) {
            const double vx = i * dv, vy = j * dv;
            const double speed = std::hypot(vx, vy);
            if (speed < lo || speed > hi) continue;
            if (count < cap) {
                vxy[2 * count] = vx;
                vxy[2 * count + 1] = vy;
            }
            ++count;
        }
    if (count > 5000000) return -1;
    return static_cast<int>(count);
}

extern "C" void ora_null_time_orders(const double* t_h, int n, int runs, uint64_t seed, double* out) {
    ora_rng r;
    ora_rng_seed(&r, seed);
    for (int k = 0; k < runs; ++k) {
        double* t = out + static_cast<size_t>(k) * n;
        std::copy(t_h, t_h + n, t);
        for (size_t i = static_cast<size_t>(n) - 1; i > 0; --i) {
            const size_t j = static_cast<size_t>(ora_rng_uniform(&r) * static_cast<double>(i + 1));
            std::swap(t[i], t[j]);
        }
    }
}

extern "C" int ora_find_peaks_rowmajor(const float* M, const int32_t* best_v, int w, int h, float floor, int radius,
                                       int edge, int cap, int32_t* px, int32_t* py, float* psnr, int32_t* pv) {
    int count = 0;
    for (int y = edge; y < h - edge; ++y)
        for (int x = edge; x < w - edge; ++x) {
            const float v = M[static_cast<size_t>(y) * w + x];
            if (v < floor) continue;
            bool peak = true;
            for (int dy = -radius; dy <= radius && peak; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= h) continue;
                for (int dx = -radius; dx <= radius; ++dx) {
                    const int xx = x + dx;
                    if ((dx == 0 && dy == 0) || xx < 0 || xx >= w) continue;
                    const float u = M[static_cast<size_t>(yy) * w + xx];
                    // a plateau yields one peak: its first pixel in row-major order
                    const bool earlier = dy < 0 || (dy == 0 && dx < 0);
                    if (u > v || (u == v && earlier)) {
                        peak = false;
                        break;
                    }
                }
            }
            if (!peak) continue;
            if (count < cap) {
                px[count] = x;
                py[count] = y;
                psnr[count] = v;
                pv[count] = best_v[static_cast<size_t>(y) * w + x];
            }
            ++count;
        }
    return count;
}

namespace {
struct PeakRec {
    int x = 0, y = 0;
    float snr = 0.0f;
    int vidx = -1;
};
}  // namespace

extern "C" int ora_sort_peaks(int count, int max_peaks, int32_t* px, int32_t* py, float* psnr, int32_t* pv) {
    std::vector<PeakRec> pk(static_cast<size_t>(count));
    for (int k = 0; k < count; ++k) pk[static_cast<size_t>(k)] = {px[k], py[k], psnr[k], pv[k]};
    std::sort(pk.begin(), pk.end(), [](const PeakRec& a, const PeakRec& b) { return a.snr > b.snr; });
    const int keep = std::min(count, max_peaks);
    for (int k = 0; k < keep; ++k) {
        px[k] = pk[static_cast<size_t>(k)].x;
        py[k] = pk[static_cast<size_t>(k)].y;
        psnr[k] = pk[static_cast<size_t>(k)].snr;
        pv[k] = pk[static_cast<size_t>(k)].vidx;
    }
    return keep;
}

namespace {

// Least squares of log N(>u) = a + k log u - inv_s2 u^2/2 on the top max_per_run*runs
// pooled peaks (search.cpp:480-515): normal equations, Gauss-Jordan with partial pivoting.
bool tail_fit_gauss(std::vector<float> pooled, double runs, double& a, double& k, double& inv_s2,
                    double max_per_run) {
    std::sort(pooled.begin(), pooled.end(), std::greater<float>());
    const size_t m = std::min(pooled.size(), static_cast<size_t>(max_per_run * runs));
    if (m < 8) return false;
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, R[3] = {0, 0, 0};
    for (size_t q = 0; q < m; ++q) {
        const double u = pooled[q];
        if (!(u > 0.0)) return false;
        const double f[3] = {1.0, std::log(u), -0.5 * u * u};
        const double yv = std::log(static_cast<double>(q + 1) / runs);
        for (int r = 0; r < 3; ++r) {
            R[r] += f[r] * yv;
            for (int c = 0; c < 3; ++c) S[r][c] += f[r] * f[c];
        }
    }
    double A[3][4];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) A[r][c] = S[r][c];
        A[r][3] = R[r];
    }
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        if (std::fabs(A[piv][c]) < 1e-300) return false;
        for (int q = 0; q < 4; ++q) std::swap(A[c][q], A[piv][q]);
        for (int r = 0; r < 3; ++r) {
            if (r == c) continue;
            const double f = A[r][c] / A[c][c];
            for (int q = c; q < 4; ++q) A[r][q] -= f * A[c][q];
        }
    }
    a = A[0][3] / A[0][0];
    k = A[1][3] / A[1][1];
    inv_s2 = A[2][3] / A[2][2];
    return std::isfinite(a) && std::isfinite(k) && inv_s2 > 0.0 && std::isfinite(inv_s2);
}

double tail_log_rate(double a, double k, double inv_s2, double u) { return a + k * std::log(u) - 0.5 * inv_s2 * u * u; }

// Bisection for the level with `rate` peaks per run on the falling branch (search.cpp:517-527).
double tail_level(double a, double k, double inv_s2, double rate) {
    const double lt = std::log(rate);
    double lo = std::max(1.0, k > 0.0 ? std::sqrt(k / inv_s2) : 1.0), hi = 100.0;
    if (tail_log_rate(a, k, inv_s2, lo) <= lt) return lo;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (tail_log_rate(a, k, inv_s2, mid) > lt) lo = mid;
        else hi = mid;
    }
    return hi;
}

// log N(>s) = a - b s between 0.67 and 30 peaks per run (search.cpp:529-553).
bool tail_fit_linear(std::vector<float> pooled, double runs, double& a, double& b) {
    std::sort(pooled.begin(), pooled.end(), std::greater<float>());
    std::vector<double> xs, ys;
    for (size_t q = 0; q < pooled.size(); ++q) {
        const double per_run = static_cast<double>(q + 1) / runs;
        if (per_run < 0.67) continue;
        if (per_run > 30.0) break;
        xs.push_back(pooled[q]);
        ys.push_back(std::log(per_run));
    }
    if (xs.size() < 8) return false;
    const double mx = std::accumulate(xs.begin(), xs.end(), 0.0) / xs.size();
    const double my = std::accumulate(ys.begin(), ys.end(), 0.0) / ys.size();
    double sxy = 0, sxx = 0;
    for (size_t q = 0; q < xs.size(); ++q) {
        sxy += (xs[q] - mx) * (ys[q] - my);
        sxx += (xs[q] - mx) * (xs[q] - mx);
    }
    if (sxx <= 0.0) return false;
    const double slope = sxy / sxx;
    if (slope >= 0.0) return false;
    b = -slope;
    a = my + b * mx;
    return true;
}

}  // namespace

extern "C" void ora_null_threshold(const float* pooled_p, int count, int runs, double target_false_per_run,
                                   double min_snr_floor, double* out) {
    std::vector<float> pooled(pooled_p, pooled_p + count);
    const double target = std::max(1e-6, target_false_per_run);
    double a = 0, b = 0, k = 0, inv_s2 = 0, thr = 0;
    int method = 2;
    if (tail_fit_gauss(pooled, runs, a, k, inv_s2, 3000.0)) {
        thr = tail_level(a, k, inv_s2, target);
        method = 0;
    } else if (tail_fit_linear(pooled, runs, a, b)) {
        thr = (a - std::log(target)) / b;
        method = 1;
    } else {
        const float mx = pooled.empty() ? 0.0f : *std::max_element(pooled.begin(), pooled.end());
        thr = mx + 0.5;
    }
    thr = std::max(thr, min_snr_floor);
    out[0] = thr;
    out[1] = a;
    out[2] = k;
    out[3] = inv_s2;
    out[4] = b;
    out[5] = method;
}
