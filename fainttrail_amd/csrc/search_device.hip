// Device-resident blind search (see search_device.hpp).
#include "search_device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <numeric>
#include <stdexcept>

#include "../../include/ftk_hip.h"
#include "stack_launch.hpp"
#include "stack_plan.hpp"

namespace ftk {

namespace {

constexpr double kPiD = 3.14159265358979323846;
constexpr double kFwhmToSigmaS = 1.0 / 2.354820045030949;

void scheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

template <typename T>
T* salloc(size_t count, const char* what) {
    T* p = nullptr;
    scheck(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(1, count) * sizeof(T)), what);
    return p;
}

// xoshiro256** seeded through splitmix64 (Blackman & Vigna; the reference's Rng,
// src/random.hpp:14-31)
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t z = seed;
        for (auto& v : s) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t q = z;
            q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ULL;
            q = (q ^ (q >> 27)) * 0x94d049bb133111ebULL;
            v = q ^ (q >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool fit_gauss(const std::vector<float>& desc, double runs, double& a, double& k, double& inv_s2,
               const double* device_sums) {
    const size_t m = std::min(desc.size(), static_cast<size_t>(3000.0 * runs));
    if (m < 8) return false;
    double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, R[3] = {0, 0, 0};
    if (device_sums) {
        if (device_sums[9] != 0.0) return false;  // a non-positive peak value
        const double* d = device_sums;
        S[0][0] = d[0]; S[0][1] = S[1][0] = d[1]; S[0][2] = S[2][0] = d[2];
        S[1][1] = d[3]; S[1][2] = S[2][1] = d[4]; S[2][2] = d[5];
        R[0] = d[6]; R[1] = d[7]; R[2] = d[8];
    } else {
        for (size_t q = 0; q < m; ++q) {
            const double u = desc[q];
            if (!(u > 0.0)) return false;
            const double f[3] = {1.0, std::log(u), -0.5 * u * u};
            const double yv = std::log(static_cast<double>(q + 1) / runs);
            for (int r = 0; r < 3; ++r) {
                R[r] += f[r] * yv;
                for (int c = 0; c < 3; ++c) S[r][c] += f[r] * f[c];
            }
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

bool fit_linear(const std::vector<float>& desc, double runs, double& a, double& b) {
    std::vector<double> xs, ys;
    for (size_t q = 0; q < desc.size(); ++q) {
        const double per = static_cast<double>(q + 1) / runs;
        if (per < 0.67) continue;
        if (per > 30.0) break;
        xs.push_back(desc[q]);
        ys.push_back(std::log(per));
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

struct HostPeak {
    int x, y;
    float snr;
    int vidx;
};

}  // namespace

std::vector<std::pair<double, double>> velocity_grid(double span_h, double vmin, double vmax, double tol, double* dv_out) {
    if (!(span_h > 0.0)) throw std::runtime_error("invalid time span");
    if (vmax <= 0.0 || vmax < vmin) throw std::runtime_error("invalid velocity range");
    const double dv = 4.0 * std::max(0.05, tol) / span_h;
    if (dv_out) *dv_out = dv;
    const int K = static_cast<int>(std::ceil(vmax / dv));
    const double lo = std::max(0.0, vmin - 0.5 * dv), hi = vmax + 0.5 * dv;
    std::vector<std::pair<double, double>> g;
    for (int j = -K; j <= K; ++j)
        for (int i = -K; i <= K; ++i) {
            const double vx = i * dv, vy = j * dv;
            const double sp = std::hypot(vx, vy);
            if (sp >= lo && sp <= hi) g.emplace_back(vx, vy);
        }
    if (g.size() > 5000000) throw std::runtime_error("velocity grid too large; narrow the range");
    return g;
}

std::vector<std::vector<double>> null_time_orders(const std::vector<double>& t_h, int runs, uint64_t seed) {
    Xoshiro rng(seed);
    std::vector<std::vector<double>> out;
    for (int k = 0; k < runs; ++k) {
        std::vector<double> t = t_h;
        for (size_t i = t.size() - 1; i > 0; --i) {
            const size_t j = static_cast<size_t>(rng.uniform() * static_cast<double>(i + 1));
            std::swap(t[i], t[j]);
        }
        out.push_back(std::move(t));
    }
    return out;
}

// set_threshold (src/search.cpp:573-612)
void set_threshold(const SearchConfig& cfg, int width, int height, double fwhm, size_t n_motions, int runs,
                   const std::vector<float>& pooled, SearchOutput& out, const double* device_sums) {
    if (cfg.threshold_snr > 0.0) {
        out.threshold = cfg.threshold_snr;
        out.method = 3;
        return;
    }
    if (runs > 0) {
        std::vector<float> desc(pooled);
        std::sort(desc.begin(), desc.end(), std::greater<float>());
        const double target = std::max(1e-6, cfg.target_false_per_run);
        double a = 0, b = 0, k = 0, inv_s2 = 0;
        if (fit_gauss(desc, runs, a, k, inv_s2, device_sums)) {
            out.null_a = a;
            out.null_k = k;
            out.null_inv_s2 = inv_s2;
            out.threshold = tail_level(a, k, inv_s2, target);
            out.method = 0;
        } else if (fit_linear(desc, runs, a, b)) {
            out.null_a = a;
            out.null_b = b;
            out.threshold = (a - std::log(target)) / b;
            out.method = 1;
        } else {
            const float mx = pooled.empty() ? 0.0f : *std::max_element(pooled.begin(), pooled.end());
            out.threshold = mx + 0.5;
            out.method = 2;
        }
        out.threshold = std::max(out.threshold, cfg.min_snr_floor);
        return;
    }
    const double resel = 4.0 * kPiD * std::pow(fwhm * kFwhmToSigmaS, 2);
    const double trials = std::max(1.0, width * height / resel * n_motions * 0.5);
    double lo = 3.0, hi = 12.0;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (trials * (0.5 * std::erfc(mid / std::sqrt(2.0))) > cfg.target_false_per_run) lo = mid;
        else hi = mid;
    }
    out.threshold = std::max(hi, cfg.min_snr_floor);
    out.method = 4;
}

DeviceSearch::DeviceSearch(int n_, int w_, int h_) : n(n_), w(w_), h(h_), stage_(n_, w_, h_) {
    const size_t plane = static_cast<size_t>(w) * h;
    sub_ = salloc<float>(plane * n, "allocate frames");
    filt_ = salloc<float>(plane * n, "allocate frames");
    max_ = salloc<float>(plane, "allocate maps");
    best_ = salloc<int>(plane, "allocate maps");
    nmax_ = salloc<float>(plane, "allocate maps");
    nbest_ = salloc<int>(plane, "allocate maps");
    pk_cap_ = std::max<size_t>(1 << 16, plane / 8);
    pk_ = salloc<int>(1 + 4 * pk_cap_, "allocate peaks");
}

DeviceSearch::~DeviceSearch() {
    for (void* p : {static_cast<void*>(sub_), static_cast<void*>(filt_), static_cast<void*>(max_),
                    static_cast<void*>(best_), static_cast<void*>(nmax_), static_cast<void*>(nbest_),
                    static_cast<void*>(pk_)})
        if (p) (void)hipFree(p);
}

// find_peaks (src/search.cpp:105-135): device local maxima, then the reference's
// host order: row-major list, std::sort by descending S/N, truncation.
PeakList DeviceSearch::peaks(const float* m, const int* b, float floor, int radius, int edge, size_t max_peaks,
                             hipStream_t s) {
    int* count = pk_;
    int* px = pk_ + 1;
    int* py = px + pk_cap_;
    float* ps = reinterpret_cast<float*>(py + pk_cap_);
    int* pv = reinterpret_cast<int*>(ps + pk_cap_);
    scheck(hipMemsetAsync(count, 0, sizeof(int), s), "clear peak count");
    scheck(launch_peaks(m, b, w, h, floor, radius, edge, static_cast<int>(pk_cap_), count, px, py, ps, pv, s), "peaks");
    int c = 0;
    scheck(hipMemcpyAsync(&c, count, sizeof(int), hipMemcpyDeviceToHost, s), "download peak count");
    scheck(hipStreamSynchronize(s), "peaks");
    if (static_cast<size_t>(c) > pk_cap_) throw std::runtime_error("too many local maxima for the peak buffer");
    std::vector<int> hx(c), hy(c), hv(c);
    std::vector<float> hs(c);
    if (c > 0) {
        scheck(hipMemcpyAsync(hx.data(), px, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
        scheck(hipMemcpyAsync(hy.data(), py, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
        scheck(hipMemcpyAsync(hs.data(), ps, sizeof(float) * c, hipMemcpyDeviceToHost, s), "download peaks");
        scheck(hipMemcpyAsync(hv.data(), pv, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
        scheck(hipStreamSynchronize(s), "peaks");
    }
    std::vector<HostPeak> pk(static_cast<size_t>(c));
    for (int q = 0; q < c; ++q) pk[static_cast<size_t>(q)] = {hx[q], hy[q], hs[q], hv[q]};
    // the device appends in any order; the reference scans row-major
    std::sort(pk.begin(), pk.end(), [](const HostPeak& a, const HostPeak& b) { return a.y != b.y ? a.y < b.y : a.x < b.x; });
    std::sort(pk.begin(), pk.end(), [](const HostPeak& a, const HostPeak& b) { return a.snr > b.snr; });
    if (pk.size() > max_peaks) pk.resize(max_peaks);
    PeakList out;
    for (const auto& p : pk) {
        out.x.push_back(p.x);
        out.y.push_back(p.y);
        out.snr.push_back(p.snr);
        out.vidx.push_back(p.vidx);
    }
    return out;
}

SearchOutput DeviceSearch::run(const float* raw, long long stride, const double* mjd, const SearchConfig& cfg,
                               float* d_max_snr, int* d_best_v, hipStream_t s) {
    if (!(cfg.fwhm_px > 0.0)) throw std::runtime_error("search: fwhm_px must be given (> 0)");
    if (n < 5) throw std::runtime_error("synthetic tracking needs at least 5 frames");
    if (w < 16 || h < 16) throw std::runtime_error("frames are too small");
    SearchOutput out;
    const size_t plane = static_cast<size_t>(w) * h;
    double t0 = wall();
    auto lap = [&](const char* name) {
        scheck(hipStreamSynchronize(s), name);
        const double t1 = wall();
        out.timings.emplace_back(name, t1 - t0);
        t0 = t1;
    };
    // times relative to the middle of the sequence (preprocess.cpp:541-553)
    double tmin = mjd[0], tmax = mjd[0];
    for (int i = 0; i < n; ++i) {
        tmin = std::min(tmin, mjd[i]);
        tmax = std::max(tmax, mjd[i]);
    }
    out.t_ref_mjd = 0.5 * (tmin + tmax);
    out.span_h = (tmax - tmin) * 24.0;
    if (out.span_h <= 0.0) throw std::runtime_error("frames must span a non-zero time interval");
    for (int i = 0; i < n; ++i) out.t_h.push_back((mjd[i] - out.t_ref_mjd) * 24.0);

    StageParams prm;
    prm.fwhm_px = cfg.fwhm_px;
    prm.clip_sigma = cfg.clip_sigma;
    prm.cr_sigma = cfg.cr_sigma;
    prm.static_exclude = cfg.static_exclude;
    prm.mask_scatter = cfg.mask_scatter;
    prm.weight_scatter = cfg.weight_scatter;
    stage_.background(raw, stride, sub_, static_cast<long long>(plane), nullptr, s);
    stage_.build(sub_, static_cast<long long>(plane), out.t_h.data(), prm, filt_, static_cast<long long>(plane),
                 nullptr, nullptr, nullptr, nullptr, s);
    lap("prepare");

    out.grid = velocity_grid(out.span_h, cfg.vmin_pxh, cfg.vmax_pxh, cfg.grid_tol_px, &out.dv);
    const int nv = static_cast<int>(out.grid.size());
    const int min_count = std::max(3, static_cast<int>(std::ceil(cfg.min_coverage * n)));
    const bool auto_thr = cfg.threshold_snr <= 0.0;
    const int runs = auto_thr ? std::max(0, cfg.null_runs) : 0;
    const int nms = std::max(2, static_cast<int>(std::lround(cfg.fwhm_px)));
    const int edge = std::max(cfg.edge_px, 1);

    std::vector<int32_t> dx(static_cast<size_t>(nv) * n), dy(dx.size());
    auto pass = [&](const std::vector<double>& t, float* m, int* b) {
        for (int j = 0; j < nv; ++j)
            for (int i = 0; i < n; ++i) {
                dx[static_cast<size_t>(j) * n + i] = static_cast<int32_t>(std::lround(out.grid[static_cast<size_t>(j)].first * t[static_cast<size_t>(i)]));
                dy[static_cast<size_t>(j) * n + i] = static_cast<int32_t>(std::lround(out.grid[static_cast<size_t>(j)].second * t[static_cast<size_t>(i)]));
            }
        ftk_options opt;
        opt.device = -1;
        opt.max_mb = 0;
        opt.launch_adds = 1e11;
        ftk_plan* plan = nullptr;
        char err[512] = {0};
        if (ftk_plan_create(n, w, h, nv, dx.data(), dy.data(), min_count, &opt, &plan, err, sizeof err) != FTK_OK)
            throw std::runtime_error(err);
        const int st = ftk_plan_run(plan, filt_, plane, m, b, s, err, sizeof err);
        // the plan's device tables must outlive the launch
        scheck(hipStreamSynchronize(s), "stack pass");
        ftk_plan_destroy(plan);
        if (st != FTK_OK) throw std::runtime_error(err);
    };

    float* m_real = d_max_snr ? d_max_snr : max_;
    int* b_real = d_best_v ? d_best_v : best_;
    pass(out.t_h, m_real, b_real);
    lap("stack (real)");
    const float floor = static_cast<float>(auto_thr ? cfg.min_snr_floor : cfg.threshold_snr);
    out.peaks = peaks(m_real, b_real, floor, nms, edge, 100000, s);
    lap("peaks (real)");

    const auto orders = null_time_orders(out.t_h, runs, cfg.seed);
    float* mn = (m_real == max_) ? nmax_ : max_;
    int* bn = (b_real == best_) ? nbest_ : best_;
    for (int k = 0; k < runs; ++k) {
        const std::vector<double>& t = orders[static_cast<size_t>(k)];
        stage_.build(sub_, static_cast<long long>(plane), t.data(), prm, filt_, static_cast<long long>(plane), nullptr,
                     nullptr, nullptr, nullptr, s);
        lap("null rebuild");
        pass(t, mn, bn);
        lap("stack (null)");
        const PeakList npk = peaks(mn, bn, 3.0f, nms, edge, 20000, s);
        out.pooled.insert(out.pooled.end(), npk.snr.begin(), npk.snr.end());
        lap("peaks (null)");
    }

    double sums[10] = {0};
    const double* dsums = nullptr;
    if (cfg.fit_on_device && runs > 0) {
        std::vector<float> desc(out.pooled);
        std::sort(desc.begin(), desc.end(), std::greater<float>());
        const int m = static_cast<int>(std::min(desc.size(), static_cast<size_t>(3000.0 * runs)));
        if (m >= 8) {
            float* dpool = salloc<float>(m, "allocate pooled peaks");
            const int blocks = (m + 255) / 256;
            double* part = salloc<double>(static_cast<size_t>(blocks) * 9, "allocate sums");
            int* bad = salloc<int>(1, "allocate flag");
            scheck(hipMemcpyAsync(dpool, desc.data(), sizeof(float) * m, hipMemcpyHostToDevice, s), "upload peaks");
            scheck(hipMemsetAsync(bad, 0, sizeof(int), s), "clear flag");
            scheck(launch_tail_sums(dpool, m, static_cast<double>(runs), part, bad, s), "tail sums");
            std::vector<double> hp(static_cast<size_t>(blocks) * 9);
            int hbad = 0;
            scheck(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s), "download sums");
            scheck(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s), "download flag");
            scheck(hipStreamSynchronize(s), "tail sums");
            for (int bl = 0; bl < blocks; ++bl)
                for (int q = 0; q < 9; ++q) sums[q] += hp[static_cast<size_t>(bl) * 9 + q];
            sums[9] = hbad ? 1.0 : 0.0;
            dsums = sums;
            (void)hipFree(dpool);
            (void)hipFree(part);
            (void)hipFree(bad);
        }
    }
    set_threshold(cfg, w, h, cfg.fwhm_px, static_cast<size_t>(nv), runs, out.pooled, out, dsums);
    // select_peaks (search.cpp:614-620)
    for (size_t q = 0; q < out.peaks.snr.size(); ++q) {
        if (!(out.peaks.snr[q] >= out.threshold) || out.peaks.vidx[q] < 0) continue;
        if (static_cast<int>(out.selected.x.size()) >= cfg.max_candidates) break;
        out.selected.x.push_back(out.peaks.x[q]);
        out.selected.y.push_back(out.peaks.y[q]);
        out.selected.snr.push_back(out.peaks.snr[q]);
        out.selected.vidx.push_back(out.peaks.vidx[q]);
    }
    lap("threshold fit");
    return out;
}

}  // namespace ftk
