// Device-resident blind search (see search_device.hpp).
#include "search_device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <exception>
#include <stdexcept>
#include <thread>

#include "../../include/ftk_hip.h"
#include "stack_launch.hpp"
#include "stack_plan.hpp"
#include "refine.hpp"
#include "screen.hpp"
#include "store.hpp"
#include "fuse.hpp"

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

double expected_null_peaks(const SearchOutput& out, double snr) {
    if (out.null_inv_s2 > 0.0 && snr > 0.0) return std::exp(tail_log_rate(out.null_a, out.null_k, out.null_inv_s2, snr));
    if (out.null_b > 0.0) return std::exp(out.null_a - out.null_b * snr);
    return 0.0;
}

int build_lanes(int n, long long plane, int passes, int asked) {
    if (passes < 2) return 1;
    if (asked > 0) return std::min({asked, 4, passes});
    const double floats = static_cast<double>(n) * static_cast<double>(plane);
    return floats <= 16777216.0 ? 4 : floats <= 67108864.0 ? 2 : 1;
}

// inverse_distance_grid (src/parallax.cpp:167-185): {0} when the near-Earth search is
// off or the largest parallax displacement at the nearest distance stays within the
// grid tolerance (straight tracks suffice); else steps + 1 equal steps from 0 to
// 1 / neo_min_dist_au, neighbours 2 tol apart at the frame of largest displacement, at
// most neo_max_trials of them.
std::vector<double> inverse_distance_grid(const SearchConfig& cfg, int n, double* dq_out) {
    if (dq_out) *dq_out = 0.0;
    if (cfg.neo_min_dist_au <= 0.0) return {0.0};
    if (static_cast<int>(cfg.par_x.size()) != n || static_cast<int>(cfg.par_y.size()) != n)
        throw std::runtime_error("the near-Earth search needs the parallax model: par_x and par_y for every frame");
    double amp = 0.0;  // parallax_amplitude_px (src/parallax.cpp:159-163)
    for (int i = 0; i < n; ++i)
        amp = std::max(amp, std::hypot(cfg.par_x[static_cast<size_t>(i)], cfg.par_y[static_cast<size_t>(i)]));
    const double qmax = 1.0 / cfg.neo_min_dist_au;
    const double tol = std::max(0.05, cfg.grid_tol_px);
    if (amp * qmax <= tol) return {0.0};
    int steps = static_cast<int>(std::ceil(amp * qmax / (2.0 * tol)));
    steps = std::min(steps, std::max(1, cfg.neo_max_trials - 1));
    std::vector<double> q(static_cast<size_t>(steps) + 1);
    for (int k = 0; k <= steps; ++k) q[static_cast<size_t>(k)] = qmax * k / steps;
    if (dq_out) *dq_out = qmax / steps;
    return q;
}

// The frame whose own time each of a pass's frame times is: the first frame with that
// time (with_times, src/preprocess.cpp:503-515, gives a null run's frame i the parallax
// of that frame). For distinct times this is the pass's frame order.
static std::vector<int> time_sources(const std::vector<double>& t_h, const std::vector<double>& times) {
    std::vector<int> src(times.size());
    for (size_t i = 0; i < times.size(); ++i) {
        const auto it = std::find(t_h.begin(), t_h.end(), times[i]);
        if (it == t_h.end()) throw std::runtime_error("null run times are not a permutation of the frame times");
        src[i] = static_cast<int>(it - t_h.begin());
    }
    return src;
}

// A pass's parallax displacements: those of the frames whose times it takes.
static void pass_parallax(const SearchConfig& cfg, const std::vector<int>& src, std::vector<double>& px,
                   std::vector<double>& py) {
    px.assign(src.size(), 0.0);
    py.assign(src.size(), 0.0);
    if (cfg.par_x.empty()) return;
    for (size_t i = 0; i < src.size(); ++i) {
        px[i] = cfg.par_x[static_cast<size_t>(src[i])];
        py[i] = cfg.par_y[static_cast<size_t>(src[i])];
    }
}

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

std::vector<std::vector<int>> null_time_perms(int n, int runs, uint64_t seed) {
    Xoshiro rng(seed);
    std::vector<std::vector<int>> out;
    for (int k = 0; k < runs; ++k) {
        std::vector<int> p(static_cast<size_t>(n));
        std::iota(p.begin(), p.end(), 0);
        for (size_t i = p.size() - 1; i > 0; --i) {
            const size_t j = static_cast<size_t>(rng.uniform() * static_cast<double>(i + 1));
            std::swap(p[i], p[j]);
        }
        out.push_back(std::move(p));
    }
    return out;
}

PeakList order_peaks(int c, const int* x, const int* y, const float* snr, const int* vidx, size_t max_peaks) {
    std::vector<HostPeak> pk(static_cast<size_t>(std::max(c, 0)));
    for (int q = 0; q < c; ++q) pk[static_cast<size_t>(q)] = {x[q], y[q], snr[q], vidx[q]};
    // the device appends in any order; the reference scans row-major
    std::sort(pk.begin(), pk.end(), [](const HostPeak& a, const HostPeak& b) { return a.y != b.y ? a.y < b.y : a.x < b.x; });
    std::sort(pk.begin(), pk.end(), [](const HostPeak& a, const HostPeak& b) { return a.snr > b.snr; });
    if (pk.size() > max_peaks) pk.resize(max_peaks);
    PeakList out;
    out.x.reserve(pk.size());
    out.y.reserve(pk.size());
    out.snr.reserve(pk.size());
    out.vidx.reserve(pk.size());
    for (const auto& p : pk) {
        out.x.push_back(p.x);
        out.y.push_back(p.y);
        out.snr.push_back(p.snr);
        out.vidx.push_back(p.vidx);
    }
    return out;
}

void PassSet::release() {
    for (void* p : {static_cast<void*>(filt), static_cast<void*>(keys), static_cast<void*>(maxm),
                    static_cast<void*>(best), static_cast<void*>(tabs), static_cast<void*>(perm),
                    static_cast<void*>(counts), static_cast<void*>(pk), static_cast<void*>(scl),
                    static_cast<void*>(maxb), static_cast<void*>(queue_ws)})
        if (p) (void)hipFree(p);
    if (h_counts) (void)hipHostFree(h_counts);
    *this = PassSet();
}

void PassSet::reserve(int cap_, long long frames_, long long frame_floats_, long long plane_, long long tab_ints_,
                      long long pk_cap_) {
    if (cap_ <= cap && frames_ == frames && frame_floats_ <= frame_floats && plane_ == plane && tab_ints_ <= tab_ints &&
        pk_cap_ <= pk_cap && filt)
        return;
    release();
    const size_t c = static_cast<size_t>(cap_);
    filt = salloc<float>(c * frames_ * frame_floats_, "allocate pass frames");
    keys = salloc<unsigned long long>(c * plane_, "allocate pass keys");
    maxm = salloc<float>(c * plane_, "allocate pass maps");
    best = salloc<int>(c * plane_, "allocate pass maps");
    tabs = salloc<int>(c * tab_ints_, "allocate pass tables");
    perm = salloc<int>(c * frames_, "allocate pass orders");
    counts = salloc<int>(c, "allocate peak counts");
    pk = salloc<int>(c * 4 * pk_cap_, "allocate pass peaks");
    scl = salloc<float>(c * frames_, "allocate frame scales");
    maxb = salloc<unsigned>(c * frames_, "allocate frame scratch");
    queue_ws = salloc<int>(64, "allocate queue counters");
    scheck(hipHostMalloc(reinterpret_cast<void**>(&h_counts), sizeof(int) * c, hipHostMallocDefault),
           "allocate pinned counts");
    cap = cap_;
    frames = frames_;
    frame_floats = frame_floats_;
    plane = plane_;
    layout = -1;
    tab_ints = tab_ints_;
    pk_cap = pk_cap_;
}

double PassSet::device_bytes() const {
    if (!filt) return 0.0;
    const double c = cap;
    return c * frames * frame_floats * 4.0 + c * plane * (8.0 + 4.0 + 4.0) + c * tab_ints * 4.0 + c * frames * 4.0 +
           c * 4.0 + c * 4.0 * pk_cap * 4.0 + c * frames * (4.0 + 4.0) + 64.0 * 4.0;
}

double DeviceSearch::device_bytes() const {
    const double plane = static_cast<double>(w) * h;
    double b = 2.0 * plane * n * 4.0 + 5.0 * plane * 4.0 + (1.0 + 4.0 * pk_cap_) * 4.0;
    b += static_cast<double>(scr_tab_cap_) * 8.0 + static_cast<double>(scr_cap_) + static_cast<double>(scr_art_cap_);
    b += static_cast<double>(tab_bytes_);
    b += stage_.device_bytes();
    for (const PassSet& set : sets_) b += set.device_bytes();
    if (real_) b += plane * n * 8.0 + plane * 4.0;
    b += static_cast<double>(ref_cap_);
    for (const BuildLane& l : lanes_) {
        if (l.stage) b += l.stage->device_bytes();
        if (l.filt) b += plane * n * 4.0;
    }
    return b;
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
    sky_ = salloc<float>(plane, "allocate maps");
    scheck(hipStreamCreateWithFlags(&build_stream_, hipStreamNonBlocking), "create stream");
    scheck(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking), "create stream");
    for (int q = 0; q < 2; ++q) {
        scheck(hipEventCreateWithFlags(&ev_built_[q], hipEventDisableTiming), "create event");
        scheck(hipEventCreateWithFlags(&ev_stacked_[q], hipEventDisableTiming), "create event");
        scheck(hipEventCreateWithFlags(&ev_counts_[q], hipEventDisableTiming), "create event");
        scheck(hipEventCreateWithFlags(&ev_ring_in_[q], hipEventDisableTiming), "create event");
        scheck(hipEventCreateWithFlags(&ev_ring_out_[q], hipEventDisableTiming), "create event");
    }
    scheck(hipStreamCreateWithFlags(&ring_stream_, hipStreamNonBlocking), "create stream");
    scheck(hipEventCreateWithFlags(&ev_start_, hipEventDisableTiming), "create event");
    scheck(hipEventCreateWithFlags(&ev_tables_, hipEventDisableTiming), "create event");
    scheck(hipEventCreateWithFlags(&ev_upload_, hipEventDisableTiming), "create event");
}

DeviceSearch::~DeviceSearch() {
    for (void* p : {static_cast<void*>(sub_), static_cast<void*>(filt_), static_cast<void*>(max_),
                    static_cast<void*>(best_), static_cast<void*>(nmax_), static_cast<void*>(nbest_),
                    static_cast<void*>(pk_), static_cast<void*>(sky_), static_cast<void*>(scr_tab_),
                    static_cast<void*>(scr_pk_), static_cast<void*>(scr_art_), static_cast<void*>(real_),
                    static_cast<void*>(res_), static_cast<void*>(rmask_), static_cast<void*>(ref_dev_)})
        if (p) (void)hipFree(p);
    for (void* p : {static_cast<void*>(scr_tab_host_), static_cast<void*>(scr_pk_host_),
                    static_cast<void*>(scr_art_host_), static_cast<void*>(ref_host_), static_cast<void*>(pk_host_)})
        if (p) (void)hipHostFree(p);
    for (auto& set : sets_) set.release();
    for (int q = 0; q < 2; ++q)
        for (hipEvent_t e : {ev_built_[q], ev_stacked_[q], ev_counts_[q], ev_ring_in_[q], ev_ring_out_[q]})
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ev_start_, ev_tables_, ev_upload_})
        if (e) (void)hipEventDestroy(e);
    if (tab_dev_) (void)hipFree(tab_dev_);
    if (tab_host_) (void)hipHostFree(tab_host_);
    if (build_stream_) (void)hipStreamDestroy(build_stream_);
    for (BuildLane& l : lanes_) {
        if (l.filt) (void)hipFree(l.filt);
        for (hipEvent_t e : l.built)
            if (e) (void)hipEventDestroy(e);
        if (l.stream) (void)hipStreamDestroy(l.stream);
        l.stage.reset();
    }
    if (copy_stream_) (void)hipStreamDestroy(copy_stream_);
    if (ring_stream_) (void)hipStreamDestroy(ring_stream_);
}

// find_peaks (src/search.cpp:105-135): device local maxima, then the reference's
// host order.
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
    return order_peaks(c, hx.data(), hy.data(), hs.data(), hv.data(), max_peaks);
}

// The run's tables of the null screen, in one pinned block and one copy on stream s:
// velocities by index, sqrt(c) for the S/N normalisation (track_snr, search.cpp:158),
// the detector test's offsets (search.cpp:217-223: the same accumulation, so the same
// doubles), the registration offsets and the frame times of every pass; for a
// near-Earth search also every pass's parallax displacements (those of the frames
// whose times it takes) and the inverse distances.
void DeviceSearch::screen_tables(const SearchConfig& cfg, const SearchOutput& out,
                                 const std::vector<std::vector<double>>& times, hipStream_t s) {
    std::vector<double> steps;
    const double span = std::max(2.0, 2.0 * cfg.fwhm_px);
    for (double d = -span; d <= span; d += 0.5) steps.push_back(d);
    const bool reg = !cfg.reg_dx.empty();
    const bool curved = out.dq > 0.0;
    const size_t nv = out.grid.size(), nq = out.inv_dist.size();
    const size_t need = 2 * nv + (n + 1) + steps.size() + (reg ? 2 * static_cast<size_t>(n) : 0) +
                        times.size() * static_cast<size_t>(n) * (curved ? 3 : 1) + nq;
    if (need > scr_tab_cap_) {
        if (scr_tab_) scheck(hipFree(scr_tab_), "free screen tables");
        if (scr_tab_host_) scheck(hipHostFree(scr_tab_host_), "free screen tables");
        scr_tab_ = scr_tab_host_ = nullptr;
        scr_tab_cap_ = 0;
        const size_t cap = need + need / 4;
        scr_tab_ = salloc<double>(cap, "allocate screen tables");
        scheck(hipHostMalloc(reinterpret_cast<void**>(&scr_tab_host_), sizeof(double) * cap, hipHostMallocDefault),
               "allocate screen tables");
        scr_tab_cap_ = cap;
    }
    double* q = scr_tab_host_;
    for (const auto& v : out.grid) {
        *q++ = v.first;
        *q++ = v.second;
    }
    for (int c = 0; c <= n; ++c) *q++ = std::sqrt(static_cast<double>(c));
    for (double d : steps) *q++ = d;
    if (reg) {
        for (double d : cfg.reg_dx) *q++ = d;
        for (double d : cfg.reg_dy) *q++ = d;
    }
    for (const auto& t : times) q = std::copy(t.begin(), t.end(), q);
    if (curved) {
        std::vector<double> px, py;
        for (const auto& t : times) {
            pass_parallax(cfg, time_sources(out.t_h, t), px, py);
            q = std::copy(px.begin(), px.end(), q);
        }
        for (const auto& t : times) {
            pass_parallax(cfg, time_sources(out.t_h, t), px, py);
            q = std::copy(py.begin(), py.end(), q);
        }
    }
    q = std::copy(out.inv_dist.begin(), out.inv_dist.end(), q);
    scheck(hipMemcpyAsync(scr_tab_, scr_tab_host_, sizeof(double) * need, hipMemcpyHostToDevice, s),
           "upload screen tables");
    scr_nv_ = static_cast<int>(nv);
    scr_nq_ = static_cast<int>(nq);
    scr_passes_ = static_cast<int>(times.size());
    scr_curved_ = curved;
    scr_nsteps_ = static_cast<int>(steps.size());
    scr_edge_r_ = static_cast<int>(std::ceil(2.0 * cfg.fwhm_px)) + 1;
    scr_reg_ = reg;
}

// null_tested_count and the artefact tests of search.cpp:773-778 for several passes
// in one upload (a frames descriptor per pass, then the peaks), one launch and one
// download; returns the flags per pass.
std::vector<std::vector<unsigned char>> DeviceSearch::screen_passes(const SearchConfig& cfg,
                                                                    const std::vector<ScreenPass>& sp, hipStream_t s,
                                                                    SearchOutput& out) {
    std::vector<std::vector<unsigned char>> flags(sp.size());
    std::vector<size_t> off(sp.size() + 1, 0);
    for (size_t k = 0; k < sp.size(); ++k) {
        const size_t c = cfg.null_tested > 0 ? std::min(sp[k].peaks->x.size(), static_cast<size_t>(cfg.null_tested)) : 0;
        off[k + 1] = off[k] + c;
        if (sp[k].storage != sp[0].storage) throw std::runtime_error("null screen: passes of one call share a storage");
    }
    const size_t total = off.back();
    if (total == 0) return flags;
    const size_t desc = sizeof(ScreenFrames) * sp.size();
    const size_t bytes = desc + sizeof(int) * 4 * total;
    if (bytes > scr_cap_ || total > scr_art_cap_) {
        for (void* p : {static_cast<void*>(scr_pk_), static_cast<void*>(scr_art_)})
            if (p) scheck(hipFree(p), "free screen buffers");
        for (void* p : {static_cast<void*>(scr_pk_host_), static_cast<void*>(scr_art_host_)})
            if (p) scheck(hipHostFree(p), "free screen buffers");
        scr_pk_ = scr_pk_host_ = nullptr;
        scr_art_ = scr_art_host_ = nullptr;
        scr_cap_ = scr_art_cap_ = 0;
        const size_t cap = std::max<size_t>(bytes + bytes / 4, 65536);
        const size_t acap = std::max<size_t>(total + total / 4, 4096);
        scr_pk_ = salloc<char>(cap, "allocate screen buffers");
        scr_art_ = salloc<unsigned char>(acap, "allocate screen buffers");
        scheck(hipHostMalloc(reinterpret_cast<void**>(&scr_pk_host_), cap, hipHostMallocDefault), "allocate screen buffers");
        scheck(hipHostMalloc(reinterpret_cast<void**>(&scr_art_host_), acap, hipHostMallocDefault),
               "allocate screen buffers");
        scr_cap_ = cap;
        scr_art_cap_ = acap;
    }
    const double* vel = scr_tab_;
    const double* sqrtc = vel + 2 * static_cast<size_t>(scr_nv_);
    const double* steps = sqrtc + (n + 1);
    const double* reg = steps + scr_nsteps_;
    const double* times = scr_reg_ ? reg + 2 * static_cast<size_t>(n) : reg;
    const size_t tn = static_cast<size_t>(scr_passes_) * n;
    const double* par_x = scr_curved_ ? times + tn : nullptr;
    const double* par_y = scr_curved_ ? times + 2 * tn : nullptr;
    const double* qs = times + (scr_curved_ ? 3 : 1) * tn;
    auto* hdesc = reinterpret_cast<ScreenFrames*>(scr_pk_host_);
    int* hint = reinterpret_cast<int*>(scr_pk_host_ + desc);
    for (size_t k = 0; k < sp.size(); ++k) {
        const size_t at = static_cast<size_t>(sp[k].pass) * n;
        hdesc[k] = ScreenFrames{sp[k].frames, sp[k].fstride, sp[k].pitch, 0, sp[k].scale, times + at,
                                par_x ? par_x + at : nullptr, par_y ? par_y + at : nullptr};
        const PeakList& pk = *sp[k].peaks;
        for (size_t j = 0; j < off[k + 1] - off[k]; ++j) {
            const int v = pk.vidx[j];
            if (v < 0 || v >= scr_nv_ * scr_nq_) throw std::runtime_error("null screen: motion index out of range");
            hint[off[k] + j] = pk.x[j];
            hint[total + off[k] + j] = pk.y[j];
            hint[2 * total + off[k] + j] = v;
            hint[3 * total + off[k] + j] = static_cast<int>(k);
        }
    }
    scheck(hipMemcpyAsync(scr_pk_, scr_pk_host_, bytes, hipMemcpyHostToDevice, s), "upload peaks");
    const int* dint = reinterpret_cast<const int*>(scr_pk_ + desc);
    ScreenArgs a;
    a.passes = reinterpret_cast<const ScreenFrames*>(scr_pk_);
    a.pslot = dint + 3 * total;
    a.storage = sp[0].storage;
    a.sub = sub_;
    a.sub_stride = static_cast<long long>(w) * h;
    a.sky = sky_;
    a.reg_dx = scr_reg_ ? reg : nullptr;
    a.reg_dy = scr_reg_ ? reg + n : nullptr;
    a.vel = vel;
    a.nv = scr_nv_;
    a.qs = scr_curved_ ? qs : nullptr;
    a.sqrtc = sqrtc;
    a.steps = steps;
    a.nsteps = scr_nsteps_;
    a.px = dint;
    a.py = dint + total;
    a.pv = dint + 2 * total;
    a.count = static_cast<int>(total);
    a.art = scr_art_;
    a.n = n;
    a.w = w;
    a.h = h;
    a.edge_r = scr_edge_r_;
    a.max_frame_fraction = cfg.max_frame_fraction;
    scheck(launch_screen(a, s), "null screen");
    scheck(hipMemcpyAsync(scr_art_host_, scr_art_, total, hipMemcpyDeviceToHost, s), "download screen flags");
    scheck(hipStreamSynchronize(s), "null screen");
    for (size_t k = 0; k < sp.size(); ++k) {
        flags[k].assign(scr_art_host_ + off[k], scr_art_host_ + off[k + 1]);
        out.screened += static_cast<int>(off[k + 1] - off[k]);
    }
    return flags;
}

void pool_null(const PeakList& pk, const std::vector<unsigned char>& art, SearchOutput& out) {
    for (size_t k = 0; k < pk.snr.size(); ++k) {
        if (k < art.size() && art[k]) {
            ++out.artefacts;
            continue;
        }
        out.pooled.push_back(pk.snr[k]);
    }
}

bool plan_stored_frames(const ftk_plan* pl, const uint16_t** origin, long long* stride, int* pitch,
                        const float** scale);

namespace {

struct Lap {
    SearchOutput& out;
    hipStream_t s;
    double t0;
    Lap(SearchOutput& o, hipStream_t st) : out(o), s(st), t0(wall()) {}
    void operator()(const char* name) {
        scheck(hipStreamSynchronize(s), name);
        const double t1 = wall();
        out.timings.emplace_back(name, t1 - t0);
        t0 = t1;
    }
};

StageParams stage_params(const SearchConfig& cfg) {
    StageParams prm;
    prm.fwhm_px = cfg.fwhm_px;
    prm.clip_sigma = cfg.clip_sigma;
    prm.cr_sigma = cfg.cr_sigma;
    prm.static_exclude = cfg.static_exclude;
    prm.mask_scatter = cfg.mask_scatter;
    prm.weight_scatter = cfg.weight_scatter;
    return prm;
}

int min_count_of(const SearchConfig& cfg, int n) { return std::max(3, static_cast<int>(std::ceil(cfg.min_coverage * n))); }

// A pass's shift table for its frame times t and parallax displacements (par_x,
// par_y; read only for q != 0): motion j = k nv + iv, frame i: lround(v t_i + q par_i)
// (run_pass's curved tables, search.cpp:73-95), which for q = 0 is the linear table
// lround(v t_i) (linear_shift_table, src/backend.cpp:114-123).
void shifts_for(const SearchOutput& out, const double* par_x, const double* par_y, const std::vector<double>& t, int n,
                std::vector<int32_t>& dx, std::vector<int32_t>& dy) {
    const size_t nv = out.grid.size(), nm = out.motions();
    dx.resize(nm * n);
    dy.resize(nm * n);
    for (size_t k = 0; k < out.inv_dist.size(); ++k) {
        const double q = out.inv_dist[k];
        for (size_t iv = 0; iv < nv; ++iv) {
            const size_t j = k * nv + iv;
            const auto& v = out.grid[iv];
            for (int i = 0; i < n; ++i) {
                const double ti = t[static_cast<size_t>(i)];
                double ddx = v.first * ti, ddy = v.second * ti;
                if (q != 0.0) {
                    ddx += q * par_x[i];
                    ddy += q * par_y[i];
                }
                dx[j * n + i] = static_cast<int32_t>(std::lround(ddx));
                dy[j * n + i] = static_cast<int32_t>(std::lround(ddy));
            }
        }
    }
}


void set_segments(StackPlan& p, int segments) {
    if (segments <= 0) return;
    const int nseg = std::max(1, std::min(segments, p.nb));
    p.blocks_per_seg = (p.nb + nseg - 1) / nseg;
    p.nseg = (p.nb + p.blocks_per_seg - 1) / p.blocks_per_seg;
}

}  // namespace

SearchOutput DeviceSearch::run(const float* raw, long long stride, const double* mjd, const SearchConfig& cfg,
                               float* d_max_snr, int* d_best_v, hipStream_t s) {
    if (!(cfg.fwhm_px > 0.0)) throw std::runtime_error("search: fwhm_px must be given (> 0)");
    if (n < 5) throw std::runtime_error("synthetic tracking needs at least 5 frames");
    if (w < 16 || h < 16) throw std::runtime_error("frames are too small");
    if (cfg.passes_per_launch < 0) throw std::runtime_error("search: passes_per_launch must be >= 0");
    if ((!cfg.reg_dx.empty() || !cfg.reg_dy.empty()) &&
        (cfg.reg_dx.size() != static_cast<size_t>(n) || cfg.reg_dy.size() != static_cast<size_t>(n)))
        throw std::runtime_error("search: registration offsets must be given for every frame (or for none)");
    SearchOutput out;
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
    out.grid = velocity_grid(out.span_h, cfg.vmin_pxh, cfg.vmax_pxh, cfg.grid_tol_px, &out.dv);
    out.inv_dist = inverse_distance_grid(cfg, n, &out.dq);
    if (out.motions() > 5000000)
        throw std::runtime_error("near-Earth search too large; narrow the velocity range or raise --neo");
    try {
        if (cfg.passes_per_launch == 1) return run_sequential(raw, stride, cfg, std::move(out), d_max_snr, d_best_v, s);
        return run_batched(raw, stride, cfg, std::move(out), d_max_snr, d_best_v, s);
    } catch (...) {
        // a run that stops early may leave work queued that reads this searcher's
        // buffers and tables: let it finish before the buffers can be reused or freed
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(build_stream_);
        (void)hipStreamSynchronize(copy_stream_);
        (void)hipStreamSynchronize(ring_stream_);
        for (BuildLane& l : lanes_) (void)hipStreamSynchronize(l.stream);
        throw;
    }
}

SearchOutput DeviceSearch::run_sequential(const float* raw, long long stride, const SearchConfig& cfg,
                                          SearchOutput out, float* d_max_snr, int* d_best_v, hipStream_t s) {
    const size_t plane = static_cast<size_t>(w) * h;
    Lap lap(out, s);
    const StageParams prm = stage_params(cfg);
    stage_.background(raw, stride, sub_, static_cast<long long>(plane), nullptr, s);
    // the candidate stage keeps the real pass's frames, residuals and mask
    if (cfg.refines()) reserve_real();
    float* const real_frames = cfg.refines() ? real_ : filt_;
    stage_.build(sub_, static_cast<long long>(plane), out.t_h.data(), prm, real_frames, static_cast<long long>(plane),
                 sky_, nullptr, cfg.refines() ? rmask_ : nullptr, nullptr, s, cfg.refines() ? res_ : nullptr);
    lap("prepare");

    const int nm = static_cast<int>(out.motions());
    const int min_count = min_count_of(cfg, n);
    const bool auto_thr = cfg.threshold_snr <= 0.0;
    const int runs = auto_thr ? std::max(0, cfg.null_runs) : 0;
    const int nms = std::max(2, static_cast<int>(std::lround(cfg.fwhm_px)));
    const int edge = std::max(cfg.edge_px, 1);
    out.passes_per_launch = 1;
    // work shape: asked, else tuned, else automatic (each plan resolves it)
    const int ask_rows = cfg.tile_rows > 0 ? cfg.tile_rows : tuned_tile_rows;
    const int ask_queue = cfg.queue >= 0 ? cfg.queue : tuned_queue;

    std::vector<int32_t> dx, dy;
    struct PlanDel {
        void operator()(ftk_plan* p) const { ftk_plan_destroy(p); }
    };
    using PlanPtr = std::unique_ptr<ftk_plan, PlanDel>;
    // one pass with a plan of its own; the plan (and its 16-bit frames) is returned
    auto pass = [&](const std::vector<double>& t, const double* par_x, const double* par_y, const float* frames,
                    float* m, int* b) -> PlanPtr {
        shifts_for(out, par_x, par_y, t, n, dx, dy);
        ftk_options opt;
        opt.device = -1;
        opt.max_mb = 0;
        opt.launch_adds = 1e12;
        opt.pad = 1;
        opt.tile_rows = ask_rows;
        opt.queue = ask_queue;
        ftk_plan* plan = nullptr;
        char err[512] = {0};
        if (ftk_plan_create(n, w, h, nm, dx.data(), dy.data(), min_count, &opt, &plan, err, sizeof err) != FTK_OK)
            throw std::runtime_error(err);
        PlanPtr held(plan);
        int st = FTK_OK;
        if (cfg.storage == kStoreF32) {
            st = ftk_plan_run(plan, frames, plane, m, b, s, err, sizeof err);
        } else {
            st = ftk_plan_store(plan, frames, plane, cfg.storage, s, err, sizeof err);
            if (st == FTK_OK) st = ftk_plan_run_stored(plan, m, b, s, err, sizeof err);
        }
        // the plan's device tables must outlive the launch
        scheck(hipStreamSynchronize(s), "stack pass");
        if (st != FTK_OK) throw std::runtime_error(err);
        return held;
    };

    float* m_real = d_max_snr ? d_max_snr : max_;
    int* b_real = d_best_v ? d_best_v : best_;
    {
        // the shape the plans resolve (the same rule for every pass of the run)
        shifts_for(out, cfg.par_x.data(), cfg.par_y.data(), out.t_h, n, dx, dy);
        const StackPlan probe = make_stack_plan(n, w, h, nm, dx.data(), dy.data(), min_count);
        const StackShape sh = resolve_shape(probe, cfg.storage, ask_rows, ask_queue, h);
        out.tile_rows = sh.tile_rows;
        out.queue = sh.queue;
    }
    pass(out.t_h, cfg.par_x.data(), cfg.par_y.data(), real_frames, m_real, b_real);
    lap("stack (real)");
    const float floor = static_cast<float>(auto_thr ? cfg.min_snr_floor : cfg.threshold_snr);
    out.peaks = peaks(m_real, b_real, floor, nms, edge, 100000, s);
    lap("peaks (real)");

    const auto orders = null_time_orders(out.t_h, runs, cfg.seed);
    if (runs > 0 && cfg.null_tested > 0) screen_tables(cfg, out, orders, s);
    float* mn = (m_real == max_) ? nmax_ : max_;
    int* bn = (b_real == best_) ? nbest_ : best_;
    for (int k = 0; k < runs; ++k) {
        const std::vector<double>& t = orders[static_cast<size_t>(k)];
        stage_.build(sub_, static_cast<long long>(plane), t.data(), prm, filt_, static_cast<long long>(plane), nullptr,
                     nullptr, nullptr, nullptr, s);
        lap("null rebuild");
        std::vector<double> ppx, ppy;
        pass_parallax(cfg, time_sources(out.t_h, t), ppx, ppy);
        const PlanPtr held = pass(t, ppx.data(), ppy.data(), filt_, mn, bn);
        lap("stack (null)");
        const PeakList npk = peaks(mn, bn, 3.0f, nms, edge, 20000, s);
        // the artefact tests read the frames the pass stacked: fp32 as built, or the
        // plan's 16-bit frames
        ScreenPass sp{filt_, static_cast<long long>(plane), w, kStoreF32, nullptr, k, &npk};
        if (cfg.storage != kStoreF32) {
            const uint16_t* origin = nullptr;
            if (!plan_stored_frames(held.get(), &origin, &sp.fstride, &sp.pitch, &sp.scale))
                throw std::runtime_error("null screen: the pass stored no 16-bit frames");
            sp.frames = origin;
            sp.storage = cfg.storage;
        }
        const auto art = screen_passes(cfg, {sp}, s, out);
        pool_null(npk, art[0], out);
        lap("peaks (null)");
    }
    finish(cfg, runs, nm, out, s);
    lap("threshold fit");
    if (cfg.refines()) {
        refine(cfg, out, s);
        lap("refine");
    }
    return out;
}

// The batched form. Pass 0 is the real pass, pass 1 + k the k-th null run. The real
// pass goes first on its own, then the null runs in groups of G, through two sets of
// device buffers: the stage builds of a group run
// on the build stream while the previous group stacks, each group's passes stack in
// one multi-pass launch (interior tiles) plus one for the ring tiles, and the peak
// lists of a group come back on the copy stream while the next group stacks. One
// plan serves every pass: the per-frame rows of its tables are permuted on the device
// for each null run's frame order. Every pass computes exactly what the one-pass form
// computes, in the same order per pixel and motion.
SearchOutput DeviceSearch::run_batched(const float* raw, long long stride, const SearchConfig& cfg, SearchOutput out,
                                       float* d_max_snr, int* d_best_v, hipStream_t s) {
    const long long plane = static_cast<long long>(w) * h;
    Lap lap(out, s);
    const StageParams prm = stage_params(cfg);
    const int nm = static_cast<int>(out.motions());
    const int min_count = min_count_of(cfg, n);
    const bool auto_thr = cfg.threshold_snr <= 0.0;
    const int runs = auto_thr ? std::max(0, cfg.null_runs) : 0;
    const int nms = std::max(2, static_cast<int>(std::lround(cfg.fwhm_px)));
    const int edge = std::max(cfg.edge_px, 1);
    const int passes = 1 + runs;
    if (cfg.refines()) reserve_real();

    // background on the caller's stream (the build lanes wait for it below)
    stage_.background(raw, stride, sub_, plane, nullptr, s);

    // the plan of the real pass, shared by all passes
    std::vector<int32_t> dx, dy;
    shifts_for(out, cfg.par_x.data(), cfg.par_y.data(), out.t_h, n, dx, dy);
    // the host plan: the previous run's when its shift tables and coverage count are the
    // same (the plan depends on nothing else; its tiles are set below), else a new one
    if (!plan_cache_ || plan_key_mc_ != min_count || plan_key_dx_ != dx || plan_key_dy_ != dy) {
        plan_cache_.reset();
        plan_cache_ = std::make_unique<StackPlan>(make_stack_plan(n, w, h, nm, dx.data(), dy.data(), min_count));
        plan_key_dx_ = dx;
        plan_key_dy_ = dy;
        plan_key_mc_ = min_count;
    }
    StackPlan& plan = *plan_cache_;
    plan.tile_rows = kTileY;
    plan_tiles(plan, h);
    // work shape: asked, else tuned, else automatic; the one-workgroup-per-item
    // launches take 16-row tiles only, and 16-bit frames need room for the plan's spread
    const StackShape shape = resolve_shape(plan, cfg.storage, cfg.tile_rows > 0 ? cfg.tile_rows : tuned_tile_rows,
                                           cfg.queue >= 0 ? cfg.queue : tuned_queue, h);
    if (shape.tile_rows != plan.tile_rows) {
        plan.tile_rows = shape.tile_rows;
        plan_tiles(plan, h);
    }
    out.tile_rows = shape.tile_rows;
    out.queue = shape.queue;
    const int segments = cfg.segments > 0 ? cfg.segments : tuned_segments;
    set_segments(plan, segments);
    out.segments = plan.nseg;
    const long long n_shift = static_cast<long long>(plan.nb) * n * 2 * kVel;
    const long long n_box = static_cast<long long>(plan.nb) * n * 4;
    const long long n_off = static_cast<long long>(plan.nb) * n * kVel;
    const long long n_lat = static_cast<long long>(plan.nb_lat) * n * 2;
    const long long tab_ints = n_shift + n_box + n_off + n_lat;

    // frames of a pass in a zero-bordered buffer that holds every staged box (the ring
    // tiles then stage like interior ones), when that costs at most twice their size;
    // 16-bit frames (with a scale per frame) always use it, with its 16-bit geometry
    const int st = cfg.storage;
    if (st != kStoreF32 && st != kStoreF16 && st != kStoreBF16) throw std::runtime_error("search: unknown frame storage");
    const PadGeometry pad32 = pad_geometry(plan, 0, h, 0, plan.tiles_x, plan.tiles_y);
    const PadGeometry pad = st == kStoreF32 ? pad32 : pad_geometry16(pad32);
    const bool use_pad = st != kStoreF32 || static_cast<double>(pad.pitch) * pad.rows <= 2.0 * static_cast<double>(plane);
    const long long fstride = use_pad ? static_cast<long long>(pad.rows) * pad.pitch : plane;  // elements
    const long long forigin = use_pad ? static_cast<long long>(pad.top) * pad.pitch + pad.left : 0;
    const int ebytes = st == kStoreF32 ? 4 : 2;
    const long long ffloats = (fstride * ebytes + 3) / 4;  // a frame's buffer in floats
    const long long layout = use_pad ? ((static_cast<long long>(pad.pitch) << 40) ^ (static_cast<long long>(pad.rows) << 20) ^
                                        (static_cast<long long>(pad.top) << 10) ^ (static_cast<long long>(st) << 8) ^ pad.left)
                                     : 0;

    // group size: requested, tuned, or all passes, within half the free device memory
    int G = cfg.passes_per_launch > 0 ? cfg.passes_per_launch : (tuned_passes > 0 ? tuned_passes : passes);
    G = std::max(1, std::min(G, passes));
    {
        size_t free_b = 0, total_b = 0;
        scheck(hipMemGetInfo(&free_b, &total_b), "memory info");
        const double per_pass = 4.0 * n * ffloats + 24.0 * plane + 4.0 * tab_ints + 16.0 * static_cast<double>(pk_cap_) +
                                12.0 * n + 2.0 * sizeof(StackArgs);
        const double held = sets_[0].filt ? 2.0 * sets_[0].cap * per_pass : 0.0;  // reusable buffers count as free
        const int fit = static_cast<int>((0.5 * (static_cast<double>(free_b) + held)) / (2.0 * per_pass));
        if (fit < 1) throw std::runtime_error("search: not enough device memory for a pass");
        G = std::min(G, fit);
    }
    out.passes_per_launch = G;
    // group k holds passes [gstart[k], gstart[k + 1]): the real pass alone first (the
    // only stage build the GPU waits for before it stacks; the next group's builds run
    // beside its launches), then groups of G. A persistent launch holds every CU until
    // it ends, so nothing runs beside it: then the real pass opens the first group.
    std::vector<int> gstart{0};
    if (passes > 1 && !shape.queue) gstart.push_back(1);
    while (gstart.back() < passes) gstart.push_back(std::min(passes, gstart.back() + G));
    const int groups = static_cast<int>(gstart.size()) - 1;
    const int nsets = groups > 1 ? 2 : 1;
    for (int q = 0; q < nsets; ++q) {
        PassSet& S = sets_[q];
        S.reserve(G, n, ffloats, plane, tab_ints, static_cast<long long>(pk_cap_));
        if (use_pad && S.layout != layout) {
            // the border must read as zeros; the interiors are rewritten by every pass
            scheck(hipMemsetAsync(S.filt, 0, sizeof(float) * S.cap * S.frames * S.frame_floats, s), "clear frames");
            S.layout = layout;
        }
    }
    // the build lanes start after the background and the border clears
    scheck(hipEventRecord(ev_start_, s), "record event");
    scheck(hipStreamWaitEvent(build_stream_, ev_start_, 0), "stream wait");

    // frame orders: identity for the real pass, the null shuffles after it; the time
    // ranks of each pass's frame times (build_stage's stable sort)
    std::vector<std::vector<int>> perms;
    perms.emplace_back(static_cast<size_t>(n));
    std::iota(perms[0].begin(), perms[0].end(), 0);
    for (auto& p : null_time_perms(n, runs, cfg.seed)) perms.push_back(std::move(p));
    // all_perm: per pass, the frame whose table row each frame takes: the first frame
    // with its time (a null run's frame i has the real pass's shifts of that frame:
    // the same time and, for curved tracks, the same parallax)
    std::vector<int> all_perm, all_rank(static_cast<size_t>(passes) * n);
    std::vector<std::vector<double>> ptimes(static_cast<size_t>(passes), std::vector<double>(static_cast<size_t>(n)));
    for (int p = 0; p < passes; ++p) {
        std::vector<double>& t = ptimes[static_cast<size_t>(p)];
        for (int i = 0; i < n; ++i)
            t[static_cast<size_t>(i)] = out.t_h[static_cast<size_t>(perms[static_cast<size_t>(p)][static_cast<size_t>(i)])];
        time_ranks(t.data(), n, all_rank.data() + static_cast<size_t>(p) * n);
        const std::vector<int> src = time_sources(out.t_h, t);
        all_perm.insert(all_perm.end(), src.begin(), src.end());
    }

    // per-run tables in one pinned block and one copy: the shared tables (the ones that
    // do not depend on the frame order), the real rows, orders, ranks and descriptors
    struct Part {
        const void* src;
        size_t bytes;
        size_t off;
    };
    std::vector<Part> parts;
    size_t total = 0;
    auto part = [&](const void* src, size_t bytes) {
        parts.push_back(Part{src, bytes, total});
        total += (bytes + 255) & ~static_cast<size_t>(255);
        return parts.size() - 1;
    };
    const size_t i_shift = part(plan.shift.data(), plan.shift.size() * 4);
    const size_t i_box = part(plan.box.data(), plan.box.size() * 4);
    const size_t i_off = part(plan.lds_off.data(), plan.lds_off.size() * 4);
    const size_t i_ubox = part(plan.ubox.data(), plan.ubox.size() * 4);
    const size_t i_motion = part(plan.motion.data(), plan.motion.size() * 4);
    const size_t i_inv = part(plan.inv_sqrt.data(), plan.inv_sqrt.size() * 4);
    const size_t i_covoff = part(plan.cov_off.data(), plan.cov_off.size() * 8);
    const size_t i_covx = part(plan.cov_x.data(), plan.cov_x.size() * 4);
    const size_t i_covy = part(plan.cov_y.data(), plan.cov_y.size() * 4);
    const size_t i_chain = part(plan.chain.data(), plan.chain.size() * 4);
    const size_t i_lat = part(plan.lat.data(), plan.lat.size() * 4);
    const size_t i_perm = part(all_perm.data(), all_perm.size() * 4);
    const size_t i_rank = part(all_rank.data(), all_rank.size() * 4);
    std::vector<StackArgs> hargs[2];
    size_t i_args[2] = {0, 0};
    for (int q = 0; q < nsets; ++q) {
        hargs[q].assign(static_cast<size_t>(kArgsKinds * G), StackArgs{});
        i_args[q] = part(hargs[q].data(), hargs[q].size() * sizeof(StackArgs));
    }
    // the previous run's copy and kernels are done with both buffers before they change
    if (tables_used_) scheck(hipEventSynchronize(ev_tables_), "previous run");
    if (total > tab_bytes_) {
        if (tab_dev_) scheck(hipFree(tab_dev_), "free tables");
        if (tab_host_) scheck(hipHostFree(tab_host_), "free tables");
        tab_dev_ = tab_host_ = nullptr;
        tab_bytes_ = 0;
        const size_t bytes = total + total / 4;
        scheck(hipMalloc(&tab_dev_, bytes), "allocate plan tables");
        scheck(hipHostMalloc(&tab_host_, bytes, hipHostMallocDefault), "allocate pinned tables");
        tab_bytes_ = bytes;
    }
    char* const dbase = static_cast<char*>(tab_dev_);
    auto dptr = [&](size_t k) { return static_cast<void*>(dbase + parts[k].off); };
    const int* d_shift0 = static_cast<const int*>(dptr(i_shift));
    const int* d_box0 = static_cast<const int*>(dptr(i_box));
    const int* d_off0 = static_cast<const int*>(dptr(i_off));
    const int* d_ubox = static_cast<const int*>(dptr(i_ubox));
    const int* d_motion = static_cast<const int*>(dptr(i_motion));
    const float* d_inv = static_cast<const float*>(dptr(i_inv));
    const long long* d_covoff = static_cast<const long long*>(dptr(i_covoff));
    const int* d_covx = static_cast<const int*>(dptr(i_covx));
    const int* d_covy = static_cast<const int*>(dptr(i_covy));
    const int* d_chain = static_cast<const int*>(dptr(i_chain));
    const int* d_lat0 = static_cast<const int*>(dptr(i_lat));
    const int* d_perm = static_cast<const int*>(dptr(i_perm));
    const int* d_rank = static_cast<const int*>(dptr(i_rank));

    // descriptors of every slot of every set: fixed for the run
    for (int q = 0; q < nsets; ++q) {
        PassSet& S = sets_[q];
        for (int g = 0; g < G; ++g) {
            int* tab = S.tabs + static_cast<long long>(g) * S.tab_ints;
            StackLaunch L{};
            if (st == kStoreF32)
                L.frames = S.filt + static_cast<long long>(g) * n * ffloats + forigin;
            else
                L.frames = reinterpret_cast<uint16_t*>(S.filt + static_cast<long long>(g) * n * ffloats) + forigin;
            L.frame_stride = ffloats * 4 / ebytes;
            L.pitch = use_pad ? pad.pitch : w;
            L.padded = use_pad;
            L.storage = st;
            L.scale = st == kStoreF32 ? nullptr : S.scl + static_cast<long long>(g) * n;
            L.row0 = 0;
            L.row1 = h;
            L.shift = tab;
            L.box = tab + n_shift;
            L.lds_off = tab + n_shift + n_box;
            L.lat = plan.nb_lat > 0 ? tab + n_shift + n_box + n_off : nullptr;
            L.ubox = d_ubox;
            L.cov_off = d_covoff;
            L.cov_x = d_covx;
            L.cov_y = d_covy;
            L.chain = d_chain;
            L.motion = d_motion;
            L.inv_sqrt = d_inv;
            L.y_begin = 0;
            L.y_end = h;
            L.tiles_x = plan.tiles_x;
            L.tiles_y = plan.tiles_y;
            L.bps = plan.blocks_per_seg;
            L.seg0 = 0;
            L.nseg = plan.nseg;
            L.keys = S.keys + static_cast<long long>(g) * plane;
            StackArgs A[kArgsKinds];
            make_stack_args(plan, L, A);
            for (int kind = 0; kind < kArgsKinds; ++kind) hargs[q][static_cast<size_t>(kind * G + g)] = A[kind];
        }
    }
    char* const hbase = static_cast<char*>(tab_host_);
    for (const Part& pt : parts)
        if (pt.bytes) std::memcpy(hbase + pt.off, pt.src, pt.bytes);
    // the copy stream is idle here (the previous run collected everything on it)
    scheck(hipMemcpyAsync(tab_dev_, tab_host_, total, hipMemcpyHostToDevice, copy_stream_), "upload plan tables");
    scheck(hipEventRecord(ev_upload_, copy_stream_), "record event");
    // the null screen runs on the copy stream, after each group's peak lists
    const bool screen = runs > 0 && cfg.null_tested > 0;
    if (screen) screen_tables(cfg, out, ptimes, copy_stream_);
    scheck(hipStreamWaitEvent(s, ev_upload_, 0), "stream wait");
    scheck(hipStreamWaitEvent(build_stream_, ev_upload_, 0), "stream wait");

    // build lanes: more for small frames, where a stage build's many small kernels
    // leave most of the GPU idle
    const int lanes = build_lanes(n, plane, passes, cfg.build_lanes);
    while (static_cast<int>(lanes_.size()) < lanes - 1) {
        BuildLane l;
        l.stage = std::make_unique<DeviceStage>(n, w, h);
        l.filt = salloc<float>(static_cast<size_t>(plane) * n, "allocate frames");
        scheck(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking), "create stream");
        for (auto& e : l.built) scheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "create event");
        lanes_.push_back(std::move(l));
    }
    // filter taps in place before the builds are queued (a copy only when they change)
    stage_.set_taps(prm, build_stream_);
    for (int l = 1; l < lanes; ++l) {
        BuildLane& L = lanes_[static_cast<size_t>(l) - 1];
        L.stage->set_taps(prm, L.stream);
        scheck(hipStreamWaitEvent(L.stream, ev_start_, 0), "stream wait");
        scheck(hipStreamWaitEvent(L.stream, ev_upload_, 0), "stream wait");
    }
    lap("plan");

    const float real_floor = static_cast<float>(auto_thr ? cfg.min_snr_floor : cfg.threshold_snr);
    std::vector<PeakList> results(static_cast<size_t>(passes));
    std::vector<std::vector<unsigned char>> arts(static_cast<size_t>(passes));

    auto enqueue = [&](int grp) {
        const int q = grp % nsets;
        PassSet& S = sets_[q];
        const int p0 = gstart[static_cast<size_t>(grp)], cnt = gstart[static_cast<size_t>(grp) + 1] - p0;
        auto build_pass = [&](int g) {
            const int p = p0 + g;
            // pass g of the group on build lane g mod lanes
            const int lane = g % lanes;
            BuildLane* const L = lane > 0 ? &lanes_[static_cast<size_t>(lane) - 1] : nullptr;
            hipStream_t bs = L ? L->stream : build_stream_;
            DeviceStage& stg = L ? *L->stage : stage_;
            float* scratch = L ? L->filt : filt_;
            int* tab = S.tabs + static_cast<long long>(g) * S.tab_ints;
            const int* perm = d_perm + static_cast<long long>(p) * n;
            const int* rk = d_rank + static_cast<long long>(p) * n;
            const PermuteTable pt[4] = {{d_shift0, tab, plan.nb, 2 * kVel},
                                        {d_box0, tab + n_shift, plan.nb, 4},
                                        {d_off0, tab + n_shift + n_box, plan.nb, kVel},
                                        {d_lat0, tab + n_shift + n_box + n_off, plan.nb_lat, 2}};
            scheck(launch_permute_rows(pt, n_lat > 0 ? 4 : 3, n, perm, bs), "permute tables");
            float* slot = S.filt + static_cast<long long>(g) * n * ffloats;
            const long long fs = ffloats * 4 / ebytes;  // frame stride in elements
            // the real pass also leaves the static sky (the same for every frame order)
            float* sky = p == 0 ? sky_ : nullptr;
            if (p == 0 && cfg.refines()) {
                // the candidate stage keeps the real pass's fp32 frames, residuals and
                // mask; the pass stacks a copy of the frames in its slot
                stg.build_ranked(sub_, plane, rk, prm, real_, plane, sky, nullptr, rmask_, nullptr, bs, res_);
                if (st != kStoreF32)
                    scheck(launch_store(real_, plane, w, n, w, h, st, reinterpret_cast<uint16_t*>(slot) + forigin, fs,
                                        pad.pitch, S.maxb + static_cast<long long>(g) * n,
                                        S.scl + static_cast<long long>(g) * n, bs),
                           "store frames");
                else if (use_pad)
                    scheck(launch_copy_rows(real_, plane, slot + forigin, fs, pad.pitch, w, h, n, bs), "pad frames");
                else
                    scheck(hipMemcpyAsync(slot, real_, sizeof(float) * n * plane, hipMemcpyDeviceToDevice, bs),
                           "copy frames");
            } else if (st != kStoreF32) {
                stg.build_ranked(sub_, plane, rk, prm, scratch, plane, sky, nullptr, nullptr,
                                 nullptr, bs);
                scheck(launch_store(scratch, plane, w, n, w, h, st, reinterpret_cast<uint16_t*>(slot) + forigin, fs, pad.pitch,
                                    S.maxb + static_cast<long long>(g) * n, S.scl + static_cast<long long>(g) * n, bs),
                       "store frames");
            } else if (use_pad) {
                // normalised while copied into the padded layout
                stg.build_ranked(sub_, plane, rk, prm, scratch, plane, sky, nullptr, nullptr,
                                 nullptr, bs, nullptr, true);
                stg.copy_scaled(scratch, plane, slot + forigin, fs, pad.pitch, bs);
            } else {
                stg.build_ranked(sub_, plane, rk, prm, slot, plane, sky, nullptr, nullptr,
                                 nullptr, bs);
            }
        };
        // Each build lane's passes are enqueued by a host thread of its own: a stage build
        // is some fifty small launches, and one thread enqueueing every lane's builds in
        // turn kept the later lanes waiting for the host. The lanes use disjoint stage
        // scratch, streams and pass slots; the device is the caller's.
        const int used = std::min(lanes, cnt);
        if (shape.queue && grp > 0) {
            // Under the persistent launch the builds wait for the previous group's
            // stacking: started beside it they get a few CUs for its whole length and slow
            // it by more than they take alone (2048^2, 256 nulls: 1.27 s per group's
            // stacking beside builds, 1.13 s alone plus 0.12 s of builds after it).
            const hipEvent_t prev = ev_stacked_[(grp - 1) % nsets];
            scheck(hipStreamWaitEvent(build_stream_, prev, 0), "stream wait");
            for (int l = 1; l < used; ++l)
                scheck(hipStreamWaitEvent(lanes_[static_cast<size_t>(l) - 1].stream, prev, 0), "stream wait");
        }
        if (used > 1) {
            int device = 0;
            scheck(hipGetDevice(&device), "current device");
            std::vector<std::exception_ptr> errs(static_cast<size_t>(used));
            std::vector<std::thread> pool;
            for (int l = 1; l < used; ++l)
                pool.emplace_back([&, l] {
                    try {
                        scheck(hipSetDevice(device), "set device");
                        for (int g = l; g < cnt; g += lanes) build_pass(g);
                    } catch (...) {
                        errs[static_cast<size_t>(l)] = std::current_exception();
                    }
                });
            try {
                for (int g = 0; g < cnt; g += lanes) build_pass(g);
            } catch (...) {
                errs[0] = std::current_exception();
            }
            for (std::thread& th : pool) th.join();
            for (const std::exception_ptr& e : errs)
                if (e) std::rethrow_exception(e);
        } else {
            for (int g = 0; g < cnt; ++g) build_pass(g);
        }
        scheck(launch_fill_keys(S.keys, cnt * plane, build_stream_), "clear keys");
        scheck(hipMemsetAsync(S.counts, 0, sizeof(int) * cnt, build_stream_), "clear peak counts");
        scheck(hipEventRecord(ev_built_[q], build_stream_), "record event");
        scheck(hipStreamWaitEvent(s, ev_built_[q], 0), "stream wait");
        for (int l = 1; l < std::min(lanes, cnt); ++l) {
            BuildLane& L = lanes_[static_cast<size_t>(l) - 1];
            scheck(hipEventRecord(L.built[q], L.stream), "record event");
            scheck(hipStreamWaitEvent(s, L.built[q], 0), "stream wait");
        }
        const StackArgs* dargs = static_cast<const StackArgs*>(dptr(i_args[q]));
        const StackArgs* hk[kArgsKinds];
        const StackArgs* dk[kArgsKinds];
        for (int kind = 0; kind < kArgsKinds; ++kind) {
            hk[kind] = hargs[q].data() + kind * G;
            dk[kind] = dargs + kind * G;
        }
        if (shape.queue) {
            // the group in one persistent launch: every kind of every pass, work items
            // from the set's counters (or, under a server, perhaps with other jobs' passes)
            if (fuser_)
                fuser_->stack(hargs[q].data(), dargs, cnt, G, S.queue_ws, s, w, h);
            else
                scheck(launch_stack_queue(hargs[q].data(), dargs, cnt, G, S.queue_ws, s), "launch stack kernel");
        } else {
            // ring tiles on their own stream, beside the interior launch
            scheck(hipEventRecord(ev_ring_in_[q], s), "record event");
            scheck(hipStreamWaitEvent(ring_stream_, ev_ring_in_[q], 0), "stream wait");
            scheck(launch_stack_multi(hk, dk, cnt, s, ring_stream_), "launch stack kernel");
            scheck(hipEventRecord(ev_ring_out_[q], ring_stream_), "record event");
            scheck(hipStreamWaitEvent(s, ev_ring_out_[q], 0), "stream wait");
        }
        scheck(launch_finalize(S.keys, cnt * plane, S.maxm, S.best, s), "finalize");
        for (int g = 0; g < cnt; ++g) {
            const int p = p0 + g;
            int* px = S.pk + static_cast<long long>(g) * 4 * S.pk_cap;
            scheck(launch_peaks(S.maxm + g * plane, S.best + g * plane, w, h, p == 0 ? real_floor : 3.0f, nms, edge,
                                static_cast<int>(S.pk_cap), S.counts + g, px, px + S.pk_cap,
                                reinterpret_cast<float*>(px + 2 * S.pk_cap), px + 3 * S.pk_cap, s),
                   "peaks");
        }
        if (p0 == 0) {
            if (d_max_snr) scheck(hipMemcpyAsync(d_max_snr, S.maxm, sizeof(float) * plane, hipMemcpyDeviceToDevice, s), "copy map");
            if (d_best_v) scheck(hipMemcpyAsync(d_best_v, S.best, sizeof(int) * plane, hipMemcpyDeviceToDevice, s), "copy map");
        }
        scheck(hipEventRecord(ev_stacked_[q], s), "record event");
        // The group's peak counts come back on the stacking stream itself: the copy stream
        // never waits for another stream's work, because streams share the 4 hardware
        // queues and such a wait would hold every stream on its queue (a build lane on the
        // copy stream's queue would not start the next group's builds before this group
        // had stacked). The host reads the counts after ev_counts_, so the peak downloads
        // and the screen on the copy stream need no device-side wait.
        scheck(hipMemcpyAsync(S.h_counts, S.counts, sizeof(int) * cnt, hipMemcpyDeviceToHost, s),
               "download peak counts");
        scheck(hipEventRecord(ev_counts_[q], s), "record event");
    };
    auto collect = [&](int grp) {
        const int q = grp % nsets;
        PassSet& S = sets_[q];
        const int p0 = gstart[static_cast<size_t>(grp)], cnt = gstart[static_cast<size_t>(grp) + 1] - p0;
        scheck(hipEventSynchronize(ev_counts_[q]), "peak counts");
        // the group's peak lists into pinned host memory (DMA copies, no host staging)
        std::vector<size_t> at(static_cast<size_t>(cnt) + 1, 0);
        for (int g = 0; g < cnt; ++g) {
            const int c = S.h_counts[g];
            if (c < 0 || c > S.pk_cap) throw std::runtime_error("too many local maxima for the peak buffer");
            at[static_cast<size_t>(g) + 1] = at[static_cast<size_t>(g)] + static_cast<size_t>(4) * c;
        }
        if (at.back() > pk_host_cap_) {
            if (pk_host_) scheck(hipHostFree(pk_host_), "free pinned peaks");
            pk_host_ = nullptr;
            pk_host_cap_ = 0;
            const size_t cap = at.back() + at.back() / 2;
            scheck(hipHostMalloc(reinterpret_cast<void**>(&pk_host_), sizeof(int) * cap, hipHostMallocDefault),
                   "allocate pinned peaks");
            pk_host_cap_ = cap;
        }
        for (int g = 0; g < cnt; ++g) {
            const int c = S.h_counts[g];
            int* v = pk_host_ + at[static_cast<size_t>(g)];
            const int* px = S.pk + static_cast<long long>(g) * 4 * S.pk_cap;
            if (c > 0)
                for (int a = 0; a < 4; ++a)
                    scheck(hipMemcpyAsync(v + static_cast<size_t>(a) * c, px + a * S.pk_cap, sizeof(int) * c,
                                          hipMemcpyDeviceToHost, copy_stream_),
                           "download peaks");
        }
        scheck(hipStreamSynchronize(copy_stream_), "download peaks");
        for (int g = 0; g < cnt; ++g) {
            const int p = p0 + g, c = S.h_counts[g];
            const int* v = pk_host_ + at[static_cast<size_t>(g)];
            results[static_cast<size_t>(p)] =
                order_peaks(c, v, v + c, reinterpret_cast<const float*>(v + 2 * c), v + 3 * c, p == 0 ? 100000 : 20000);
        }
        if (!screen) return;
        // the group's null runs through the artefact tests on their frames in this set,
        // before the set takes another group
        std::vector<ScreenPass> sp;
        for (int g = 0; g < cnt; ++g) {
            const int p = p0 + g;
            if (p == 0) continue;
            float* slot = S.filt + static_cast<long long>(g) * n * ffloats;
            const void* fr = st == kStoreF32 ? static_cast<const void*>(slot + forigin)
                                             : static_cast<const void*>(reinterpret_cast<uint16_t*>(slot) + forigin);
            sp.push_back(ScreenPass{fr, ffloats * 4 / ebytes, use_pad ? pad.pitch : w, st,
                                    st == kStoreF32 ? nullptr : S.scl + static_cast<long long>(g) * n, p,
                                    &results[static_cast<size_t>(p)]});
        }
        const auto fl = screen_passes(cfg, sp, copy_stream_, out);
        for (size_t k = 0; k < sp.size(); ++k) arts[static_cast<size_t>(sp[k].pass)] = fl[k];
    };
    for (int grp = 0; grp < groups; ++grp) {
        enqueue(grp);
        if (grp > 0) collect(grp - 1);
    }
    collect(groups - 1);
    scheck(hipStreamSynchronize(build_stream_), "stage builds");
    for (int l = 1; l < lanes; ++l) scheck(hipStreamSynchronize(lanes_[static_cast<size_t>(l) - 1].stream), "stage builds");
    scheck(hipEventRecord(ev_tables_, s), "record event");
    tables_used_ = true;
    lap("passes (batched)");

    out.peaks = std::move(results[0]);
    for (int p = 1; p < passes; ++p) pool_null(results[static_cast<size_t>(p)], arts[static_cast<size_t>(p)], out);
    finish(cfg, runs, nm, out, s);
    lap("threshold fit");
    if (cfg.refines()) {
        refine(cfg, out, s);
        lap("refine");
    }
    return out;
}

// The null-tail threshold, with the FP64 tail-fit sums on the GPU for 16-bit storage
// (or as asked), and select_peaks (search.cpp:614-620).
void DeviceSearch::finish(const SearchConfig& cfg, int runs, int motions, SearchOutput& out, hipStream_t s) {
    const bool gpu_fit = cfg.fit_on_device > 0 || (cfg.fit_on_device < 0 && cfg.storage != kStoreF32);
    fit_threshold(cfg, runs, motions, gpu_fit, out, s);
    for (size_t q = 0; q < out.peaks.snr.size(); ++q) {
        if (!(out.peaks.snr[q] >= out.threshold) || out.peaks.vidx[q] < 0) continue;
        if (static_cast<int>(out.selected.x.size()) >= cfg.max_candidates) break;
        out.selected.x.push_back(out.peaks.x[q]);
        out.selected.y.push_back(out.peaks.y[q]);
        out.selected.snr.push_back(out.peaks.snr[q]);
        out.selected.vidx.push_back(out.peaks.vidx[q]);
    }
}

double DeviceSearch::refit(const SearchConfig& cfg, const SearchOutput& last, bool on_device, int* method,
                           hipStream_t s) {
    SearchOutput o;
    o.pooled = last.pooled;
    const int runs = cfg.threshold_snr <= 0.0 ? std::max(0, cfg.null_runs) : 0;
    fit_threshold(cfg, runs, static_cast<int>(last.motions()), on_device, o, s);
    if (method) *method = o.method;
    return o.threshold;
}

void DeviceSearch::fit_threshold(const SearchConfig& cfg, int runs, int motions, bool on_device, SearchOutput& out,
                                 hipStream_t s) {
    double sums[10] = {0};
    const double* dsums = nullptr;
    if (on_device && runs > 0) {
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
    set_threshold(cfg, w, h, cfg.fwhm_px, static_cast<size_t>(motions), runs, out.pooled, out, dsums);
}

void DeviceSearch::reserve_real() {
    if (real_) return;
    const size_t plane = static_cast<size_t>(w) * h;
    real_ = salloc<float>(plane * n, "allocate candidate stage");
    res_ = salloc<float>(plane * n, "allocate candidate stage");
    rmask_ = salloc<float>(plane, "allocate candidate stage");
}

// refine (search.cpp:246-395) of the selected peaks: one pinned block of tables (frame
// times, sqrt(c), detector offsets, registration offsets, photometry weights, the
// peaks' grid positions and motions) in one copy, one launch, the figures and per-frame
// samples back in one copy; finish_candidates on the host.
void DeviceSearch::refine(const SearchConfig& cfg, SearchOutput& out, hipStream_t s) {
    const size_t nc = out.selected.x.size();
    if (nc == 0) return;
    const long long plane = static_cast<long long>(w) * h;
    const double s_psf = cfg.fwhm_px * kFwhmToSigmaS;
    const int R = static_cast<int>(std::ceil(3.0 * s_psf));  // stack_cutout's half size
    const int size = 2 * R + 1;
    std::vector<double> steps;
    const double span = std::max(2.0, 2.0 * cfg.fwhm_px);
    for (double d = -span; d <= span; d += 0.5) steps.push_back(d);
    const bool reg = !cfg.reg_dx.empty();
    const bool curved = out.dq > 0.0;
    const size_t n_tab = static_cast<size_t>(n) + (n + 1) + steps.size() + (reg ? 2 * static_cast<size_t>(n) : 0) +
                         static_cast<size_t>(size) * size + 5 * nc + (curved ? 2 * static_cast<size_t>(n) : 0);
    const size_t n_fig = nc * kRefOut;
    const size_t bytes = 8 * (n_tab + n_fig) + 4 * nc * static_cast<size_t>(n);
    if (bytes > ref_cap_) {
        if (ref_dev_) scheck(hipFree(ref_dev_), "free candidate buffers");
        if (ref_host_) scheck(hipHostFree(ref_host_), "free candidate buffers");
        ref_dev_ = ref_host_ = nullptr;
        ref_cap_ = 0;
        const size_t cap = std::max<size_t>(bytes + bytes / 4, 65536);
        ref_dev_ = salloc<char>(cap, "allocate candidate buffers");
        scheck(hipHostMalloc(reinterpret_cast<void**>(&ref_host_), cap, hipHostMallocDefault),
               "allocate candidate buffers");
        ref_cap_ = cap;
    }
    double* q = reinterpret_cast<double*>(ref_host_);
    for (double t : out.t_h) *q++ = t;
    for (int c = 0; c <= n; ++c) *q++ = std::sqrt(static_cast<double>(c));
    for (double d : steps) *q++ = d;
    if (reg) {
        for (double d : cfg.reg_dx) *q++ = d;
        for (double d : cfg.reg_dy) *q++ = d;
    }
    // photometry weights exp(-r^2 / 2 s^2) (search.cpp:367-377), with the host's exp
    for (int y = 0; y < size; ++y)
        for (int x = 0; x < size; ++x)
            *q++ = std::exp(-0.5 * ((x - R) * (x - R) + (y - R) * (y - R)) / (s_psf * s_psf));
    // per candidate: grid position, then velocity and inverse distance of motion
    // j = k nv + iv (search.cpp:732-733)
    const int nv = static_cast<int>(out.grid.size()), nm = static_cast<int>(out.motions());
    for (size_t k = 0; k < nc; ++k) {
        const int v = out.selected.vidx[k];
        if (v < 0 || v >= nm) throw std::runtime_error("refine: motion index out of range");
        *q++ = out.selected.x[k];
        *q++ = out.selected.y[k];
        *q++ = out.grid[static_cast<size_t>(v % nv)].first;
        *q++ = out.grid[static_cast<size_t>(v % nv)].second;
        *q++ = out.inv_dist[static_cast<size_t>(v / nv)];
    }
    if (curved) {
        for (double d : cfg.par_x) *q++ = d;
        for (double d : cfg.par_y) *q++ = d;
    }
    scheck(hipMemcpyAsync(ref_dev_, ref_host_, 8 * n_tab, hipMemcpyHostToDevice, s), "upload candidate tables");
    const double* d = reinterpret_cast<const double*>(ref_dev_);
    RefineArgs a;
    a.frames = real_;
    a.fstride = plane;
    a.resid = res_;
    a.sub = sub_;
    a.sky = sky_;
    a.mask = rmask_;
    a.t = d;
    a.sqrtc = d + n;
    a.steps = a.sqrtc + (n + 1);
    a.nsteps = static_cast<int>(steps.size());
    a.reg_dx = reg ? a.steps + steps.size() : nullptr;
    a.reg_dy = reg ? a.reg_dx + n : nullptr;
    a.kern = a.steps + steps.size() + (reg ? 2 * static_cast<size_t>(n) : 0);
    a.R = R;
    a.cand = a.kern + static_cast<size_t>(size) * size;
    a.count = static_cast<int>(nc);
    if (curved) {
        a.par_x = a.cand + 5 * nc;
        a.par_y = a.par_x + n;
        a.dq = out.dq;
    }
    // the two local-search levels (search.cpp:258-261)
    const double dv = out.dv;
    const double levels[2][4] = {{dv / 4, dv, 0.5, 1.0}, {dv / 8, dv / 4, 0.25, 0.5}};
    for (int l = 0; l < 2; ++l)
        a.lev[l] = RefineLevel{levels[l][0], levels[l][2], static_cast<int>(std::lround(levels[l][1] / levels[l][0])),
                               static_cast<int>(std::lround(levels[l][3] / levels[l][2]))};
    a.dv = dv;
    a.n = n;
    a.w = w;
    a.h = h;
    a.edge_r = static_cast<int>(std::ceil(2.0 * cfg.fwhm_px)) + 1;
    a.out = reinterpret_cast<double*>(ref_dev_ + 8 * n_tab);
    a.contrib = reinterpret_cast<float*>(ref_dev_ + 8 * (n_tab + n_fig));
    scheck(launch_refine(a, s), "refine");
    scheck(hipMemcpyAsync(ref_host_ + 8 * n_tab, ref_dev_ + 8 * n_tab, 8 * n_fig + 4 * nc * static_cast<size_t>(n),
                          hipMemcpyDeviceToHost, s),
           "download candidates");
    scheck(hipStreamSynchronize(s), "refine");
    out.candidates = finish_candidates(cfg, out, reinterpret_cast<const double*>(ref_host_ + 8 * n_tab),
                                       reinterpret_cast<const float*>(ref_host_ + 8 * (n_tab + n_fig)));
}

double DeviceSearch::tune(const float* raw, long long stride, const double* mjd, const SearchConfig& cfg_in, int reps,
                          hipStream_t s) {
    SearchConfig cfg = cfg_in;
    cfg.candidates = 0;  // the choice concerns the passes only
    const int runs = cfg.threshold_snr <= 0.0 ? std::max(0, cfg.null_runs) : 0;
    const int passes = 1 + runs;
    // group sizes: every pass at once, about half (two groups overlap builds with
    // stacking), and a third; segment counts around the planner's choice
    std::vector<int> gs{passes};
    for (int d : {2, 3})
        if ((passes + d - 1) / d > 1) gs.push_back((passes + d - 1) / d);
    std::sort(gs.begin(), gs.end());
    gs.erase(std::unique(gs.begin(), gs.end()), gs.end());
    cfg.passes_per_launch = passes;
    cfg.segments = 0;
    const int ask_rows = cfg.tile_rows, ask_queue = cfg.queue;  // a fixed shape is not tuned
    tuned_passes = 0;
    tuned_segments = 0;
    tuned_tile_rows = 0;
    tuned_queue = -1;
    const SearchOutput probe = run(raw, stride, mjd, cfg, nullptr, nullptr, s);
    const int base = std::max(1, probe.segments);
    std::vector<int> ss{base};
    for (int v : {base / 4, base / 2, base * 2})
        if (v >= 1) ss.push_back(v);
    std::sort(ss.begin(), ss.end());
    ss.erase(std::unique(ss.begin(), ss.end()), ss.end());
    double best = 1e300;
    int bg = passes, bs = base;
    for (int g : gs)
        for (int seg : ss) {
            cfg.passes_per_launch = g;
            cfg.segments = seg;
            double t = 1e300;
            int used = seg;
            for (int r = 0; r < std::max(1, reps); ++r) {
                const double t0 = wall();
                used = run(raw, stride, mjd, cfg, nullptr, nullptr, s).segments;
                t = std::min(t, wall() - t0);
            }
            if (t < best) {
                best = t;
                bg = g;
                bs = used;
            }
        }
    // then the launch forms and tile heights at the best group size and segment count
    // (a form or height the caller fixed stays fixed)
    int bt = 0, bq = -1;
    cfg.passes_per_launch = bg;
    cfg.segments = bs;
    const StackShape cands[] = {{16, 0}, {8, 1}, {16, 1}, {32, 1}};
    for (const StackShape& c : cands) {
        if ((ask_rows > 0 && c.tile_rows != ask_rows) || (ask_queue >= 0 && c.queue != ask_queue)) continue;
        cfg.tile_rows = c.tile_rows;
        cfg.queue = c.queue;
        double t = 1e300;
        SearchOutput o;
        for (int r = 0; r < std::max(1, reps); ++r) {
            const double t0 = wall();
            o = run(raw, stride, mjd, cfg, nullptr, nullptr, s);
            t = std::min(t, wall() - t0);
        }
        // a height the 16-bit staging cannot hold ran at another: not this candidate
        if (o.tile_rows == c.tile_rows && o.queue == c.queue && t < best) {
            best = t;
            bt = c.tile_rows;
            bq = c.queue;
        }
    }
    tuned_passes = bg;
    tuned_segments = bs;
    tuned_tile_rows = bt;
    tuned_queue = bq;
    return best;
}

}  // namespace ftk
