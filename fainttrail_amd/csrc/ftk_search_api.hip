// C-ABI of the device-resident search (include/ftk_hip.h, ftk_searcher_*).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/ftk_hip.h"
#include "search_device.hpp"
#include "stack_plan.hpp"
#include "serve.hpp"

struct ftk_searcher {
    ftk::DeviceSearch* ds = nullptr;
    int device = 0;
    ftk::SearchOutput last;
};

namespace {
void put_err2(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    const size_t k = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = '\0';
}
}  // namespace

extern "C" void ftk_search_params_default(ftk_search_params* p) {
    if (!p) return;
    const ftk::SearchConfig c;
    p->vmin_pxh = c.vmin_pxh;
    p->vmax_pxh = c.vmax_pxh;
    p->grid_tol_px = c.grid_tol_px;
    p->fwhm_px = 3.0;
    p->clip_sigma = c.clip_sigma;
    p->cr_sigma = c.cr_sigma;
    p->static_exclude = c.static_exclude;
    p->mask_scatter = c.mask_scatter;
    p->weight_scatter = c.weight_scatter;
    p->min_coverage = c.min_coverage;
    p->threshold_snr = c.threshold_snr;
    p->min_snr_floor = c.min_snr_floor;
    p->null_runs = c.null_runs;
    p->target_false_per_run = c.target_false_per_run;
    p->max_candidates = c.max_candidates;
    p->edge_px = c.edge_px;
    p->seed = c.seed;
    p->fit_on_device = c.fit_on_device;
    p->passes_per_launch = 0;
    p->segments = 0;
    p->storage = FTK_STORE_F32;
    p->null_tested = c.null_tested;
    p->max_frame_fraction = c.max_frame_fraction;
    p->reg_dx = nullptr;
    p->reg_dy = nullptr;
    p->candidates = c.candidates;
    p->half_tolerance = c.half_tolerance;
    p->tile_rows = 0;
    p->queue = c.queue;
    p->build_lanes = 0;
    p->neo_min_dist_au = c.neo_min_dist_au;
    p->neo_max_trials = c.neo_max_trials;
    p->par_x = nullptr;
    p->par_y = nullptr;
}

namespace {
ftk::SearchConfig config_of(const ftk_search_params* p) {
    ftk::SearchConfig c;
    c.vmin_pxh = p->vmin_pxh;
    c.vmax_pxh = p->vmax_pxh;
    c.grid_tol_px = p->grid_tol_px;
    c.fwhm_px = p->fwhm_px;
    c.clip_sigma = p->clip_sigma;
    c.cr_sigma = p->cr_sigma;
    c.static_exclude = p->static_exclude;
    c.mask_scatter = p->mask_scatter;
    c.weight_scatter = p->weight_scatter;
    c.min_coverage = p->min_coverage;
    c.threshold_snr = p->threshold_snr;
    c.min_snr_floor = p->min_snr_floor;
    c.null_runs = p->null_runs;
    c.target_false_per_run = p->target_false_per_run;
    c.max_candidates = p->max_candidates;
    c.edge_px = p->edge_px;
    c.seed = p->seed;
    c.fit_on_device = p->fit_on_device;
    c.passes_per_launch = p->passes_per_launch;
    c.segments = p->segments;
    c.storage = p->storage;
    c.null_tested = p->null_tested;
    c.max_frame_fraction = p->max_frame_fraction;
    c.candidates = p->candidates;
    c.half_tolerance = p->half_tolerance;
    c.tile_rows = p->tile_rows;
    c.queue = p->queue;
    c.build_lanes = p->build_lanes;
    c.neo_min_dist_au = p->neo_min_dist_au;
    c.neo_max_trials = p->neo_max_trials;
    return c;
}

// The registration offsets and parallax displacements belong to the frames: n of each,
// copied (a queued job may outlive the caller's arrays).
void set_offsets(ftk::SearchConfig& c, const ftk_search_params* p, int n) {
    if (!p->par_x != !p->par_y) throw std::runtime_error("search: give both parallax displacements (par_x, par_y) or none");
    if (p->par_x) {
        c.par_x.assign(p->par_x, p->par_x + n);
        c.par_y.assign(p->par_y, p->par_y + n);
    }
    if (!p->reg_dx && !p->reg_dy) return;
    if (!p->reg_dx || !p->reg_dy) throw std::runtime_error("search: give both registration offsets (reg_dx, reg_dy) or none");
    c.reg_dx.assign(p->reg_dx, p->reg_dx + n);
    c.reg_dy.assign(p->reg_dy, p->reg_dy + n);
}

void fill_summary(const ftk::SearchOutput& o, ftk_search_summary* sum) {
    std::memset(sum, 0, sizeof *sum);
    sum->threshold_snr = o.threshold;
    sum->method = o.method;
    sum->null_a = o.null_a;
    sum->null_k = o.null_k;
    sum->null_inv_s2 = o.null_inv_s2;
    sum->null_b = o.null_b;
    sum->dv_pxh = o.dv;
    sum->t_ref_mjd = o.t_ref_mjd;
    sum->span_h = o.span_h;
    sum->n_velocities = static_cast<int>(o.grid.size());
    sum->n_peaks = static_cast<int>(o.peaks.x.size());
    sum->n_selected = static_cast<int>(o.selected.x.size());
    sum->n_pooled = static_cast<int>(o.pooled.size());
    sum->passes_per_launch = o.passes_per_launch;
    sum->segments = o.segments;
    sum->n_screened = o.screened;
    sum->n_artefacts = o.artefacts;
    sum->n_candidates = static_cast<int>(o.candidates.size());
    sum->tile_rows = o.tile_rows;
    sum->queue = o.queue;
    sum->n_distances = static_cast<int>(o.inv_dist.size());
    sum->dq_inv_au = o.dq;
    for (const auto& t : o.timings) {
        if (t.first == "prepare") sum->t_prepare += t.second;
        else if (t.first == "stack (real)") sum->t_stack_real += t.second;
        else if (t.first == "null rebuild") sum->t_null_rebuild += t.second;
        else if (t.first == "stack (null)") sum->t_stack_null += t.second;
        else if (t.first.rfind("peaks", 0) == 0) sum->t_peaks += t.second;
        else if (t.first == "plan") sum->t_plan += t.second;
        else if (t.first == "passes (batched)") sum->t_passes += t.second;
        else if (t.first == "refine") sum->t_refine += t.second;
        else sum->t_threshold += t.second;
    }
}

void check_run_args(const ftk_searcher* s, const float* d_raw, size_t stride, const double* mjd,
                    const ftk_search_params* p) {
    if (!s || !d_raw || !mjd || !p) throw std::runtime_error("ftk_searcher_run: null argument");
    if (stride < static_cast<size_t>(s->ds->w) * s->ds->h) throw std::runtime_error("ftk_searcher_run: frame stride smaller than a frame");
    for (int i = 1; i < s->ds->n; ++i)
        if (mjd[i] < mjd[i - 1]) throw std::runtime_error("ftk_searcher_run: frames must be sorted by time");
    if (p->passes_per_launch < 0 || p->segments < 0)
        throw std::runtime_error("ftk_searcher_run: passes_per_launch and segments must be >= 0");
    if (p->tile_rows != 0 && !ftk::tile_rows_ok(p->tile_rows))
        throw std::runtime_error("ftk_searcher_run: tile_rows must be 0, 8, 16 or 32");
    if (p->queue < -1 || p->queue > 1) throw std::runtime_error("ftk_searcher_run: queue must be -1, 0 or 1");
    if (p->build_lanes < 0 || p->build_lanes > 4) throw std::runtime_error("ftk_searcher_run: build_lanes must be 0-4");
    if (p->queue == 0 && p->tile_rows != 0 && p->tile_rows != 16)
        throw std::runtime_error("ftk_searcher_run: tile_rows other than 16 need queue 1 (or -1)");
    if (p->storage != FTK_STORE_F32 && p->storage != FTK_STORE_F16 && p->storage != FTK_STORE_BF16)
        throw std::runtime_error("ftk_searcher_run: unknown frame storage");
    if (p->null_tested < 0) throw std::runtime_error("ftk_searcher_run: null_tested must be >= 0");
    if (p->candidates != 0 && p->candidates != 1) throw std::runtime_error("ftk_searcher_run: candidates must be 0 or 1");
    if (p->fit_on_device < -1 || p->fit_on_device > 1) throw std::runtime_error("ftk_searcher_run: fit_on_device must be -1, 0 or 1");
    if (!(p->neo_min_dist_au >= 0.0)) throw std::runtime_error("ftk_searcher_run: neo_min_dist_au must be >= 0");
    if (p->neo_min_dist_au > 0.0 && (!p->par_x || !p->par_y))
        throw std::runtime_error("ftk_searcher_run: the near-Earth search needs par_x and par_y for every frame");
}
}  // namespace

extern "C" int ftk_searcher_create(int n, int w, int h, int device, ftk_searcher** out, char* err, size_t err_len) {
    if (!out) return FTK_ERROR;
    *out = nullptr;
    try {
        if (device < 0) device = 0;
        if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("HIP device " + std::to_string(device) + " not present");
        auto* s = new ftk_searcher();
        s->device = device;
        try {
            s->ds = new ftk::DeviceSearch(n, w, h);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" void ftk_searcher_destroy(ftk_searcher* s) {
    if (!s) return;
    delete s->ds;
    delete s;
}

extern "C" int ftk_searcher_run(ftk_searcher* s, const float* d_raw, size_t stride, const double* mjd,
                                const ftk_search_params* p, ftk_search_summary* sum, float* d_max_snr,
                                int32_t* d_best_v, void* stream, char* err, size_t err_len) {
    try {
        check_run_args(s, d_raw, stride, mjd, p);
        hipSetDevice(s->device);
        ftk::SearchConfig c = config_of(p);
        set_offsets(c, p, s->ds->n);
        s->last = s->ds->run(d_raw, static_cast<long long>(stride), mjd, c, d_max_snr, d_best_v,
                             static_cast<hipStream_t>(stream));
        if (sum) fill_summary(s->last, sum);
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_searcher_tune(ftk_searcher* s, const float* d_raw, size_t stride, const double* mjd,
                                 const ftk_search_params* p, int reps, int* passes_per_launch, int* segments,
                                 int* tile_rows, int* queue, double* seconds, void* stream, char* err,
                                 size_t err_len) {
    try {
        check_run_args(s, d_raw, stride, mjd, p);
        hipSetDevice(s->device);
        ftk::SearchConfig c = config_of(p);
        set_offsets(c, p, s->ds->n);
        const double t = s->ds->tune(d_raw, static_cast<long long>(stride), mjd, c, reps,
                                     static_cast<hipStream_t>(stream));
        if (passes_per_launch) *passes_per_launch = s->ds->tuned_passes;
        if (segments) *segments = s->ds->tuned_segments;
        if (tile_rows) *tile_rows = s->ds->tuned_tile_rows;
        if (queue) *queue = s->ds->tuned_queue;
        if (seconds) *seconds = t;
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_searcher_refit(ftk_searcher* s, const ftk_search_params* p, int on_device, double* threshold,
                                  int* method, void* stream, char* err, size_t err_len) {
    try {
        if (!s || !p) throw std::runtime_error("ftk_searcher_refit: null argument");
        hipSetDevice(s->device);
        int m = 0;
        const double t = s->ds->refit(config_of(p), s->last, on_device != 0, &m, static_cast<hipStream_t>(stream));
        if (threshold) *threshold = t;
        if (method) *method = m;
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

namespace {
int copy_peaks(const ftk::SearchOutput& o, int which, int cap, int32_t* x, int32_t* y, float* snr, int32_t* vidx) {
    if (which == 2) {
        const int c = static_cast<int>(o.pooled.size());
        for (int k = 0; k < c && k < cap; ++k)
            if (snr) snr[k] = o.pooled[static_cast<size_t>(k)];
        return c;
    }
    const ftk::PeakList& p = which == 0 ? o.peaks : o.selected;
    const int c = static_cast<int>(p.x.size());
    for (int k = 0; k < c && k < cap; ++k) {
        if (x) x[k] = p.x[static_cast<size_t>(k)];
        if (y) y[k] = p.y[static_cast<size_t>(k)];
        if (snr) snr[k] = p.snr[static_cast<size_t>(k)];
        if (vidx) vidx[k] = p.vidx[static_cast<size_t>(k)];
    }
    return c;
}
}  // namespace

extern "C" int ftk_searcher_peaks(const ftk_searcher* s, int which, int cap, int32_t* x, int32_t* y, float* snr,
                                  int32_t* vidx) {
    if (!s) return -1;
    return copy_peaks(s->last, which, cap, x, y, snr, vidx);
}

namespace {
int copy_candidates(const ftk::SearchOutput& o, int cap, ftk_candidate* out) {
    const int c = static_cast<int>(o.candidates.size());
    for (int k = 0; k < c && k < cap && out; ++k) {
        const ftk::SearchCandidate& s = o.candidates[static_cast<size_t>(k)];
        ftk_candidate& r = out[k];
        std::memset(&r, 0, sizeof r);
        r.id = s.id;
        r.ok = s.ok ? 1 : 0;
        r.x = s.x;
        r.y = s.y;
        r.vx_pxh = s.vx_pxh;
        r.vy_pxh = s.vy_pxh;
        r.pos_err_px = s.pos_err_px;
        r.vel_err_pxh = s.vel_err_pxh;
        r.snr = s.snr;
        r.snr_coarse = s.snr_coarse;
        r.snr_half1 = s.snr_half1;
        r.snr_half2 = s.snr_half2;
        r.max_frame_fraction = s.max_frame_fraction;
        r.false_alarms_per_run = s.false_alarms_per_run;
        r.flux = s.flux;
        r.mag_inst = s.mag_inst;
        r.inv_dist_au = s.inv_dist_au;
        r.inv_dist_err_au = s.inv_dist_err_au;
        std::strncpy(r.flags, s.flags.c_str(), sizeof(r.flags) - 1);
    }
    return c;
}
}  // namespace

extern "C" int ftk_searcher_candidates(const ftk_searcher* s, int cap, ftk_candidate* out) {
    if (!s) return -1;
    return copy_candidates(s->last, cap, out);
}

// find_peaks (src/search.cpp:105-135) over device maps: the local maxima are found on
// the device (peaks_kernel), then put in the reference's order on the host. Two peaks
// never lie within each other's (2 radius + 1)^2 window (the larger one, or the earlier
// of two equal ones, excludes the other), so there are at most
// ceil(w / (radius + 1)) * ceil(h / (radius + 1)) of them: the size of the scratch list.
extern "C" int ftk_find_peaks(const float* d_max_snr, const int32_t* d_best_v, int w, int h, float floor, int radius,
                              int edge, int max_peaks, int cap, int32_t* x, int32_t* y, float* snr, int32_t* vidx,
                              int* count, void* stream, char* err, size_t err_len) {
    int* buf = nullptr;
    try {
        if (!d_max_snr || !d_best_v || !count) throw std::runtime_error("ftk_find_peaks: null argument");
        if (w < 1 || h < 1 || radius < 0 || edge < 0 || max_peaks < 0 || cap < 0)
            throw std::runtime_error("ftk_find_peaks: bad sizes");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const long long most = static_cast<long long>((w + radius) / (radius + 1)) * ((h + radius) / (radius + 1));
        if (most > (1ll << 28)) throw std::runtime_error("ftk_find_peaks: map too large");
        const size_t words = 1 + 4 * static_cast<size_t>(most);
        if (hipMalloc(reinterpret_cast<void**>(&buf), words * sizeof(int)) != hipSuccess)
            throw std::runtime_error("ftk_find_peaks: allocate peak list");
        int* px = buf + 1;
        int* py = px + most;
        float* ps = reinterpret_cast<float*>(py + most);
        int* pv = reinterpret_cast<int*>(ps + most);
        auto chk = [](hipError_t e, const char* what) {
            if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
        };
        chk(hipMemsetAsync(buf, 0, sizeof(int), s), "clear peak count");
        chk(ftk::launch_peaks(d_max_snr, d_best_v, w, h, floor, radius, edge, static_cast<int>(most), buf, px, py, ps,
                              pv, s),
            "peaks kernel");
        int c = 0;
        chk(hipMemcpyAsync(&c, buf, sizeof(int), hipMemcpyDeviceToHost, s), "download peak count");
        chk(hipStreamSynchronize(s), "peaks kernel");
        if (c > most) throw std::runtime_error("ftk_find_peaks: more local maxima than the bound");
        std::vector<int> hx(c), hy(c), hv(c);
        std::vector<float> hs(c);
        if (c > 0) {
            chk(hipMemcpyAsync(hx.data(), px, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
            chk(hipMemcpyAsync(hy.data(), py, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
            chk(hipMemcpyAsync(hs.data(), ps, sizeof(float) * c, hipMemcpyDeviceToHost, s), "download peaks");
            chk(hipMemcpyAsync(hv.data(), pv, sizeof(int) * c, hipMemcpyDeviceToHost, s), "download peaks");
            chk(hipStreamSynchronize(s), "download peaks");
        }
        (void)hipFree(buf);
        buf = nullptr;
        const ftk::PeakList pk =
            ftk::order_peaks(c, hx.data(), hy.data(), hs.data(), hv.data(), static_cast<size_t>(max_peaks));
        const int total = static_cast<int>(pk.x.size());
        for (int k = 0; k < total && k < cap; ++k) {
            if (x) x[k] = pk.x[static_cast<size_t>(k)];
            if (y) y[k] = pk.y[static_cast<size_t>(k)];
            if (snr) snr[k] = pk.snr[static_cast<size_t>(k)];
            if (vidx) vidx[k] = pk.vidx[static_cast<size_t>(k)];
        }
        *count = total;
        return FTK_OK;
    } catch (const std::exception& e) {
        if (buf) (void)hipFree(buf);
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

// ------------------------------------------------------------------ server

struct ftk_server {
    ftk::SearchServer* srv = nullptr;
};

extern "C" int ftk_server_create(int device, double budget_mb, int slots, ftk_server** out, char* err,
                                 size_t err_len) {
    if (!out) return FTK_ERROR;
    *out = nullptr;
    try {
        if (device < 0) device = 0;
        auto* s = new ftk_server();
        try {
            s->srv = new ftk::SearchServer(device, budget_mb > 0 ? budget_mb * 1048576.0 : 0.0, slots);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" void ftk_server_destroy(ftk_server* s) {
    if (!s) return;
    delete s->srv;
    delete s;
}

extern "C" int ftk_server_submit(ftk_server* s, const float* d_raw, size_t stride, int n, int w, int h,
                                 const double* mjd, const ftk_search_params* p, long long* job, char* err,
                                 size_t err_len) {
    try {
        if (!s || !p || !job) throw std::runtime_error("ftk_server_submit: null argument");
        if (p->passes_per_launch < 0 || p->segments < 0) throw std::runtime_error("ftk_server_submit: negative batching");
        if (p->storage != FTK_STORE_F32 && p->storage != FTK_STORE_F16 && p->storage != FTK_STORE_BF16)
            throw std::runtime_error("ftk_server_submit: unknown frame storage");
        if (!(p->fwhm_px > 0.0)) throw std::runtime_error("search: fwhm_px must be given (> 0)");
        if (p->null_tested < 0) throw std::runtime_error("ftk_server_submit: null_tested must be >= 0");
        if (p->candidates != 0 && p->candidates != 1) throw std::runtime_error("ftk_server_submit: candidates must be 0 or 1");
        if (p->fit_on_device < -1 || p->fit_on_device > 1) throw std::runtime_error("ftk_server_submit: fit_on_device must be -1, 0 or 1");
        ftk::SearchConfig c = config_of(p);
        set_offsets(c, p, n);
        *job = s->srv->submit(d_raw, static_cast<long long>(stride), n, w, h, mjd, std::move(c));
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_server_wait(ftk_server* s, long long job, double timeout_s, ftk_search_summary* sum,
                               double* times, char* err, size_t err_len) {
    try {
        if (!s) throw std::runtime_error("ftk_server_wait: null server");
        if (!s->srv->wait(job, timeout_s)) return FTK_TIMEOUT;
        const ftk::ServerJob* j = s->srv->job(job);
        if (times) {
            times[0] = j->t_submit;
            times[1] = j->t_admit;
            times[2] = j->t_end;
            times[3] = j->reserved;
            times[4] = j->held;
        }
        if (j->status != 2) throw std::runtime_error(j->error);
        if (sum) fill_summary(j->out, sum);
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err2(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_server_peaks(ftk_server* s, long long job, int which, int cap, int32_t* x, int32_t* y, float* snr,
                                int32_t* vidx) {
    if (!s) return -1;
    const ftk::ServerJob* j = s->srv->job(job);
    if (!j || j->status != 2) return -1;
    return copy_peaks(j->out, which, cap, x, y, snr, vidx);
}

extern "C" int ftk_server_candidates(ftk_server* s, long long job, int cap, ftk_candidate* out) {
    if (!s) return -1;
    const ftk::ServerJob* j = s->srv->job(job);
    if (!j || j->status != 2) return -1;
    return copy_candidates(j->out, cap, out);
}

extern "C" void ftk_server_release(ftk_server* s, long long job) {
    if (s) s->srv->release(job);
}

extern "C" int ftk_server_stats(ftk_server* s, double* stats) {
    if (!s || !stats) return FTK_ERROR;
    const auto st = s->srv->stats();
    const double v[13] = {st.used, st.peak, st.budget, st.idle_bytes, static_cast<double>(st.running),
                          static_cast<double>(st.queued), static_cast<double>(st.completed),
                          static_cast<double>(st.failed), static_cast<double>(st.evicted),
                          static_cast<double>(st.max_running), static_cast<double>(st.stack_launches),
                          static_cast<double>(st.fused_launches), static_cast<double>(st.fused_passes)};
    for (int k = 0; k < 13; ++k) stats[k] = v[k];
    return FTK_OK;
}

extern "C" int ftk_server_fusion(ftk_server* s, int on, double window_ms) {
    if (!s || window_ms < 0) return FTK_ERROR;
    s->srv->set_fusion(on != 0, 1e-3 * window_ms);
    return FTK_OK;
}
