// CPU oracle: the time-order-dependent preprocessing stage (TEST INFRASTRUCTURE ONLY;
// see oracle.h). Restates src/preprocess.cpp:116-486 (matched_filter,
// static_sky_column, build_stage for the untiled case), :233-254 (subtract_and_grow),
// :528-585 (per-frame background) and src/common.hpp:64-100 (median_inplace,
// robust_sigma, sample_pixels), keeping every float/double conversion where the
// reference has it. Order statistics are taken from fully sorted copies, which gives
// the same values as the reference's nth_element/max_element selection.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

constexpr double kFwhmToSigma = 1.0 / 2.354820045030949;  // psf.hpp:17

// Median of finite values; mean (in double) of the two middle values for even sizes
// (common.hpp:64-74). Empty -> 0.
double median_of(std::vector<float> v) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t mid = v.size() / 2;
    double m = v[mid];
    if (v.size() % 2 == 0) m = 0.5 * (m + static_cast<double>(v[mid - 1]));
    return m;
}

// MAD * 1.4826; deviations rounded to float before the second median (common.hpp:77-82).
double robust_sigma_of(std::vector<float> v) {
    if (v.size() < 3) return 0.0;
    const double med = median_of(v);
    for (auto& x : v) x = static_cast<float>(std::fabs(x - med));
    return 1.4826 * median_of(std::move(v));
}

// Regular sample grid of finite (and unmasked) values (common.hpp:85-100).
std::vector<float> grid_samples(const float* img, int w, int h, size_t max_samples, const float* mask, int border) {
    std::vector<float> out;
    const size_t usable = static_cast<size_t>(std::max(0, w - 2 * border)) * static_cast<size_t>(std::max(0, h - 2 * border));
    if (usable == 0) return out;
    const int step = std::max(1, static_cast<int>(std::sqrt(static_cast<double>(usable) / max_samples)));
    for (int y = border; y < h - border; y += step)
        for (int x = border; x < w - border; x += step) {
            const size_t k = static_cast<size_t>(y) * w + x;
            if (mask && mask[k] == 0.0f) continue;
            if (std::isfinite(img[k])) out.push_back(img[k]);
        }
    return out;
}

// One pixel column: global median of the finite (value, time rank) pairs and every
// frame's residual against the median that leaves out frames within `exclude` ranks
// (preprocess.cpp:154-204).
float sky_column(int n, const float* vals, const int* rank, int exclude, float* residual,
                 std::vector<std::pair<float, int>>& s) {
    s.clear();
    for (int i = 0; i < n; ++i)
        if (std::isfinite(vals[i])) s.emplace_back(vals[i], rank[i]);
    if (s.size() < std::max<size_t>(5, static_cast<size_t>(0.6 * n))) {
        for (int i = 0; i < n; ++i) residual[i] = NAN;
        return NAN;
    }
    std::sort(s.begin(), s.end());
    const size_t m_all = s.size();
    // the two middle values are added in float, the half taken in double
    const double global = 0.5 * (s[(m_all - 1) / 2].first + s[m_all / 2].first);
    for (int i = 0; i < n; ++i) {
        const float v = vals[i];
        if (!std::isfinite(v)) {
            residual[i] = NAN;
            continue;
        }
        double st = global;
        if (exclude > 0) {
            const int ri = rank[i];
            size_t m = 0;
            for (const auto& e : s)
                if (std::abs(e.second - ri) > exclude) ++m;
            if (m >= 5) {
                const size_t lo = (m - 1) / 2, hi = m / 2;
                double a = 0.0, b = 0.0;
                size_t k = 0;
                for (const auto& e : s) {
                    if (std::abs(e.second - ri) <= exclude) continue;
                    if (k == lo) a = e.first;
                    if (k == hi) {
                        b = e.first;
                        break;
                    }
                    ++k;
                }
                st = 0.5 * (a + b);  // double addition here
            }
        }
        residual[i] = static_cast<float>(v - st);
    }
    return static_cast<float>(global);
}

// Separable Gaussian matched filter, truncated at the image edges, mul then add in
// float (preprocess.cpp:116-148).
void matched_filter(const float* in, float* out, int w, int h, double sigma) {
    const int R = std::max(1, static_cast<int>(std::ceil(3.0 * sigma)));
    std::vector<float> k(static_cast<size_t>(2 * R + 1));
    double k2 = 0.0;
    for (int i = -R; i <= R; ++i) {
        const double v = std::exp(-0.5 * i * i / (sigma * sigma));
        k[static_cast<size_t>(i + R)] = static_cast<float>(v);
        k2 += v * v;
    }
    const float norm = static_cast<float>(1.0 / k2);
    std::vector<float> tmp(static_cast<size_t>(w) * h);
    for (int y = 0; y < h; ++y) {
        const float* row = in + static_cast<size_t>(y) * w;
        for (int x = 0; x < w; ++x) {
            const int a = std::max(0, x - R), b = std::min(w - 1, x + R);
            float acc = 0.0f;
            for (int j = a; j <= b; ++j) acc += row[j] * k[static_cast<size_t>(j - x + R)];
            tmp[static_cast<size_t>(y) * w + x] = acc;
        }
    }
    for (size_t q = 0; q < static_cast<size_t>(w) * h; ++q) out[q] = 0.0f;
    for (int y = 0; y < h; ++y) {
        const int a = std::max(0, y - R), b = std::min(h - 1, y + R);
        float* o = out + static_cast<size_t>(y) * w;
        for (int j = a; j <= b; ++j) {
            const float kj = k[static_cast<size_t>(j - y + R)] * norm;
            const float* t = &tmp[static_cast<size_t>(j) * w];
            for (int x = 0; x < w; ++x) o[x] += kj * t[x];
        }
    }
}

}  // namespace

extern "C" int ora_build_stage(int n, int w, int h, const float* sub, const double* t_h, const ora_stage_params* p,
                               float* filtered, float* static_sky, double* sigma_out, float* mask_out,
                               int32_t* cr_out) {
    if (n < 1 || w < 3 || h < 3 || !p) return 1;
    const size_t plane = static_cast<size_t>(w) * h;
    // time rank of each frame: stable sort by time (preprocess.cpp:319-331)
    std::vector<int> order(static_cast<size_t>(n)), rank(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t_h[a] < t_h[b]; });
    for (int r = 0; r < n; ++r) rank[static_cast<size_t>(order[static_cast<size_t>(r)])] = r;
    const int exclude = p->static_exclude >= 0 ? p->static_exclude : std::max(1, static_cast<int>(std::lround(n / 8.0)));

    std::vector<float> residual(static_cast<size_t>(n) * plane);
    std::vector<float> sky(plane);
#pragma omp parallel
    {
        std::vector<std::pair<float, int>> s;
        std::vector<float> vals(static_cast<size_t>(n)), res(static_cast<size_t>(n));
#pragma omp for schedule(static)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const size_t q = static_cast<size_t>(y) * w + x;
                for (int i = 0; i < n; ++i) vals[static_cast<size_t>(i)] = sub[static_cast<size_t>(i) * plane + q];
                sky[q] = sky_column(n, vals.data(), rank.data(), exclude, res.data(), s);
                for (int i = 0; i < n; ++i) residual[static_cast<size_t>(i) * plane + q] = res[static_cast<size_t>(i)];
            }
    }
    // per-frame residual noise (preprocess.cpp:354-365, 265-268)
    std::vector<double> sig(static_cast<size_t>(n));
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i) {
        const double s = robust_sigma_of(grid_samples(&residual[static_cast<size_t>(i) * plane], w, h, 200000, nullptr, 2));
        sig[static_cast<size_t>(i)] = s > 0.0 ? s : 1.0;
    }
    // temporal scatter -> weights and mask (preprocess.cpp:371-400)
    std::vector<float> scatter(plane);
#pragma omp parallel
    {
        std::vector<float> col;
#pragma omp for schedule(static)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const size_t q = static_cast<size_t>(y) * w + x;
                col.clear();
                for (int i = 0; i < n; ++i) {
                    const float v = residual[static_cast<size_t>(i) * plane + q];
                    if (std::isfinite(v)) col.push_back(static_cast<float>(std::fabs(v) / sig[static_cast<size_t>(i)]));
                }
                scatter[q] = col.size() < static_cast<size_t>(n) / 2 ? 1e9f : static_cast<float>(1.4826 * median_of(col));
            }
    }
    std::vector<float> mask(plane, 1.0f);
    const int dil = std::max(1, static_cast<int>(std::lround(0.5 * p->fwhm_px)));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            if (scatter[static_cast<size_t>(y) * w + x] <= p->mask_scatter) continue;
            for (int yy = std::max(0, y - dil); yy <= std::min(h - 1, y + dil); ++yy)
                for (int xx = std::max(0, x - dil); xx <= std::min(w - 1, x + dil); ++xx)
                    mask[static_cast<size_t>(yy) * w + xx] = 0.0f;
        }
    // normalise, weight, cosmic rays, clip, matched filter, renormalise (preprocess.cpp:402-481)
    const double psf_sigma = p->fwhm_px * kFwhmToSigma;
    const float clip = static_cast<float>(p->clip_sigma);
    const float wscat = static_cast<float>(p->weight_scatter);
    const float ratio = static_cast<float>(0.5 * std::exp(-0.5 / (psf_sigma * psf_sigma)));
    const float hit = static_cast<float>(p->cr_sigma);
    const int fborder = static_cast<int>(std::ceil(3 * psf_sigma)) + 1;
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i) {
        const float* r = &residual[static_cast<size_t>(i) * plane];
        std::vector<float> z(plane, 0.0f);
        const float inv = static_cast<float>(1.0 / sig[static_cast<size_t>(i)]);
        for (size_t q = 0; q < plane; ++q) {
            const float v = r[q];
            if (mask[q] == 0.0f || !std::isfinite(v)) continue;
            const float e = scatter[q] > wscat ? scatter[q] : 1.0f;
            z[q] = v * inv / e;
        }
        int cr = 0;
        std::vector<size_t> kill;
        for (int y = 1; y < h - 1; ++y)
            for (int x = 1; x < w - 1; ++x) {
                const size_t q = static_cast<size_t>(y) * w + x;
                const float v = z[q];
                if (v < hit) continue;
                const float n4 = 0.25f * (z[q - 1] + z[q + 1] + z[q - w] + z[q + w]);
                if (n4 >= ratio * v) continue;
                ++cr;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const size_t u = static_cast<size_t>(y + dy) * w + (x + dx);
                        if ((dx == 0 && dy == 0) || z[u] > 2.5f) kill.push_back(u);
                    }
            }
        for (size_t u : kill) z[u] = 0.0f;
        for (auto& v : z) v = std::clamp(v, -clip, clip);
        float* g = filtered + static_cast<size_t>(i) * plane;
        matched_filter(z.data(), g, w, h, psf_sigma);
        const double sg = robust_sigma_of(grid_samples(g, w, h, 200000, mask.data(), fborder));
        if (sg > 0.0) {
            const float ig = static_cast<float>(1.0 / sg);
            for (size_t q = 0; q < plane; ++q) g[q] *= ig;
        }
        if (cr_out) cr_out[i] = cr;
    }
    if (static_sky) std::copy(sky.begin(), sky.end(), static_sky);
    if (sigma_out) std::copy(sig.begin(), sig.end(), sigma_out);
    if (mask_out) std::copy(mask.begin(), mask.end(), mask_out);
    return 0;
}

extern "C" int ora_background(int n, int w, int h, const float* frames, float* sub, double* background, int threads) {
    (void)threads;
    const size_t plane = static_cast<size_t>(w) * h;
    const int grow = 3;  // kBlankGrow (internal.hpp:63)
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i) {
        const float* f = frames + static_cast<size_t>(i) * plane;
        float* s = sub + static_cast<size_t>(i) * plane;
        const double bg = median_of(grid_samples(f, w, h, 200000, nullptr, 0));
        background[i] = bg;
        bool any = false;
        std::vector<unsigned char> bad(plane);
        for (size_t q = 0; q < plane; ++q) {
            s[q] = std::isfinite(f[q]) ? static_cast<float>(f[q] - bg) : NAN;
            bad[q] = !std::isfinite(s[q]);
            any = any || bad[q];
        }
        if (!any) continue;
        // blank (non-finite) areas grow by `grow` pixels in a square
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                if (!bad[static_cast<size_t>(y) * w + x]) continue;
                for (int yy = std::max(0, y - grow); yy <= std::min(h - 1, y + grow); ++yy)
                    for (int xx = std::max(0, x - grow); xx <= std::min(w - 1, x + grow); ++xx)
                        s[static_cast<size_t>(yy) * w + xx] = NAN;
            }
    }
    return 0;
}
