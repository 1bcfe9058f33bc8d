// Host tail of the candidate stage: refine's closing figures and flags
// (src/search.cpp:319-391) from what the device measured, and finish_candidates
// (:627-675): duplicates and ghosts, final flags, false-alarm rates, order and ids.
// Plain C++ over a few hundred candidates; every expression keeps the reference's
// operation order, so each field is the reference's.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "refine.hpp"
#include "search_device.hpp"

namespace ftk {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kFwhmToSigma = 1.0 / 2.354820045030949;

// is_rejecting and finalise_flags (search.cpp:397-413)
bool is_rejecting(const std::string& f) {
    return f == "half" || f == "single" || f == "ghost" || f == "edge" || f == "detector";
}

void finalise_flags(SearchCandidate& c) {
    c.ok = true;
    size_t start = 0;
    while (start <= c.flags.size() && !c.flags.empty()) {
        const size_t end = c.flags.find(',', start);
        if (is_rejecting(c.flags.substr(start, end == std::string::npos ? std::string::npos : end - start)))
            c.ok = false;
        if (end == std::string::npos) break;
        start = end + 1;
    }
}

// add_flag (search.cpp:415-418): a flag once, comma separated
void add_flag(SearchCandidate& c, const std::string& f) {
    if (c.flags.find(f) != std::string::npos) return;
    c.flags += c.flags.empty() ? f : "," + f;
}

struct Refined {
    SearchCandidate c;
    std::vector<double> contrib;  // per-frame track samples (NaN where not finite)
};

}  // namespace

std::vector<SearchCandidate> finish_candidates(const SearchConfig& cfg, const SearchOutput& out, const double* fig,
                                               const float* contrib) {
    const int n = static_cast<int>(out.t_h.size());
    const size_t nc = out.selected.x.size();
    const double s_psf = cfg.fwhm_px * kFwhmToSigma;
    std::vector<Refined> rc(nc);
    for (size_t k = 0; k < nc; ++k) {
        const double* o = fig + k * kRefOut;
        SearchCandidate& c = rc[k].c;
        c.snr_coarse = out.selected.snr[k];
        c.x = o[kRefX];
        c.y = o[kRefY];
        c.vx_pxh = o[kRefVx];
        c.vy_pxh = o[kRefVy];
        c.snr = o[kRefSnr];
        c.inv_dist_au = o[kRefQ];  // 0 on straight tracks
        {
            // centroid precision of a matched-filter fit, with the empirical factor 1.25
            constexpr double kEmpirical = 1.25;
            c.pos_err_px = c.snr > 0.0 ? kEmpirical * std::sqrt(2.0) * s_psf / c.snr : 0.0;
            double t2 = 0.0;
            for (double t : out.t_h) t2 += t * t;
            t2 /= std::max(1, n);
            c.vel_err_pxh = t2 > 0.0 ? c.pos_err_px / std::sqrt(t2) : 0.0;
            if (out.dq > 0.0) {
                // likewise for the inverse distance: the parallax curve takes the place of t
                double p2 = 0.0;
                for (int i = 0; i < n; ++i)
                    p2 += 0.5 * (cfg.par_x[static_cast<size_t>(i)] * cfg.par_x[static_cast<size_t>(i)] +
                                 cfg.par_y[static_cast<size_t>(i)] * cfg.par_y[static_cast<size_t>(i)]);
                p2 /= std::max(1, n);
                c.inv_dist_err_au = p2 > 0.0 ? c.pos_err_px / std::sqrt(p2) : 0.0;
            }
        }
        c.snr_half1 = o[kRefHalf1];
        c.snr_half2 = o[kRefHalf2];
        std::vector<double>& pf = rc[k].contrib;
        pf.resize(static_cast<size_t>(n));
        double pos_sum = 0.0, maxc = 0.0;
        for (int i = 0; i < n; ++i) {
            const double v = contrib[k * static_cast<size_t>(n) + static_cast<size_t>(i)];
            pf[static_cast<size_t>(i)] = v;
            if (!std::isfinite(v)) continue;
            pos_sum += std::max(0.0, v);
            maxc = std::max(maxc, v);
        }
        c.max_frame_fraction = pos_sum > 0.0 ? maxc / pos_sum : 1.0;
        const double num = o[kRefNum], den = o[kRefDen];
        c.flux = den > 0.0 ? num / den * 2.0 * kPi * s_psf * s_psf : 0.0;
        c.mag_inst = c.flux > 0.0 ? 25.0 - 2.5 * std::log10(c.flux) : 99.0;

        std::vector<std::string> flags;
        const double expect_half = c.snr_coarse / std::sqrt(2.0);
        const double half_min = std::max(2.0, expect_half - cfg.half_tolerance);
        if (c.snr_half1 < half_min || c.snr_half2 < half_min) flags.push_back("half");
        if (n >= 8 && c.max_frame_fraction > cfg.max_frame_fraction) flags.push_back("single");
        const int inside = static_cast<int>(o[kRefInside]), masked = static_cast<int>(o[kRefMasked]);
        if (inside > 0 && masked > inside / 2) flags.push_back("masked");
        if (o[kRefEdge] != 0.0) flags.push_back("edge");
        if (o[kRefDetector] != 0.0) flags.push_back("detector");
        if (std::hypot(c.vx_pxh, c.vy_pxh) > cfg.vmax_pxh - 0.5 * out.dv) flags.push_back("vedge");
        if (std::hypot(c.vx_pxh, c.vy_pxh) < 2.0 * cfg.fwhm_px / out.span_h) flags.push_back("slow");
        for (size_t i = 0; i < flags.size(); ++i) c.flags += (i ? "," : "") + flags[i];
    }

    // track_pos (src/internal.hpp:107-116) of a candidate's track in frame f
    auto track_pos = [&](int f, const SearchCandidate& c, double& ox, double& oy) {
        const double t = out.t_h[static_cast<size_t>(f)];
        ox = c.x + c.vx_pxh * t;
        oy = c.y + c.vy_pxh * t;
        if (c.inv_dist_au != 0.0) {
            ox += c.inv_dist_au * cfg.par_x[static_cast<size_t>(f)];
            oy += c.inv_dist_au * cfg.par_y[static_cast<size_t>(f)];
        }
    };
    // duplicates and ghosts: against every stronger candidate kept so far
    std::sort(rc.begin(), rc.end(), [](const Refined& a, const Refined& b) { return a.c.snr > b.c.snr; });
    const double r_overlap = 1.5 * cfg.fwhm_px;
    std::vector<bool> drop(rc.size(), false);
    for (size_t j = 0; j < rc.size(); ++j) {
        double total = 0.0;
        for (double v : rc[j].contrib)
            if (std::isfinite(v) && v > 0.0) total += v;
        std::vector<bool> shared(static_cast<size_t>(n), false);
        int overlap_max = 0;
        for (size_t i = 0; i < j && !drop[j]; ++i) {
            if (drop[i]) continue;
            int overlap = 0;
            for (int f = 0; f < n; ++f) {
                double xi, yi, xj, yj;
                track_pos(f, rc[i].c, xi, yi);
                track_pos(f, rc[j].c, xj, yj);
                const double dx = xi - xj, dy = yi - yj;
                if (dx * dx + dy * dy < r_overlap * r_overlap) {
                    ++overlap;
                    shared[static_cast<size_t>(f)] = true;
                }
            }
            overlap_max = std::max(overlap_max, overlap);
            if (overlap >= n / 2) drop[j] = true;
        }
        if (drop[j] || overlap_max == 0) continue;
        double from_shared = 0.0;
        for (int f = 0; f < n; ++f) {
            const double v = rc[j].contrib[static_cast<size_t>(f)];
            if (shared[static_cast<size_t>(f)] && std::isfinite(v) && v > 0.0) from_shared += v;
        }
        if (total > 0.0 && from_shared / total > 0.3) add_flag(rc[j].c, "ghost");
    }
    std::vector<SearchCandidate> res;
    for (size_t j = 0; j < rc.size(); ++j) {
        if (drop[j]) continue;
        SearchCandidate c = rc[j].c;
        finalise_flags(c);
        c.false_alarms_per_run = expected_null_peaks(out, c.snr_coarse);
        res.push_back(c);
    }
    std::stable_sort(res.begin(), res.end(), [](const SearchCandidate& a, const SearchCandidate& b) {
        if (a.ok != b.ok) return a.ok;
        return a.snr > b.snr;
    });
    for (size_t i = 0; i < res.size(); ++i) res[i].id = static_cast<int>(i + 1);
    return res;
}

}  // namespace ftk
