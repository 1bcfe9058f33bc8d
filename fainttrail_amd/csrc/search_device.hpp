// Device-resident blind search: the reference's search() loop (src/search.cpp:716-809)
// with every per-pixel step on the GPU. Frames are uploaded once; the real pass and
// each time-shuffled null run rebuild the preprocessing stage and stack on the device;
// only peak lists (a few thousand values) come back to the host, which sorts them
// with libstdc++'s std::sort exactly as the reference does and fits the null tail.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "prep.hpp"
#include "stack_plan.hpp"

namespace ftk {

hipError_t launch_peaks(const float* m, const int* best_v, int w, int h, float floor, int radius, int edge, int cap,
                        int* count, int* px, int* py, float* ps, int* pv, hipStream_t s);
hipError_t launch_tail_sums(const float* pooled_desc, int m, double runs, double* partial, int* bad, hipStream_t s);
// A per-frame plan table to permute: nb blocks of n frame rows of `row` ints.
struct PermuteTable {
    const int* src;
    int* dst;
    int nb, row;
};
// Row i of every block of each table's dst = row perm[i] of its src (up to 4 tables, one launch).
hipError_t launch_permute_rows(const PermuteTable* tables, int count, int n, const int* perm, hipStream_t s);

// SearchParams fields on the path (include/fainttrail/search.hpp:18-90).
struct SearchConfig {
    double vmin_pxh = 4.0, vmax_pxh = 40.0, grid_tol_px = 0.5;
    double fwhm_px = 0.0;  // must be given (> 0): the PSF estimate runs once, outside this path
    double clip_sigma = 8.0, cr_sigma = 5.0;
    int static_exclude = -1;
    double mask_scatter = 2.5, weight_scatter = 1.5, min_coverage = 0.7;
    double threshold_snr = 0.0, min_snr_floor = 5.0;
    int null_runs = 3;
    double target_false_per_run = 0.02;
    int max_candidates = 300;
    int edge_px = 4;
    uint64_t seed = 12345;
    // where the FP64 tail-fit sums run: 1 on the GPU (tolerance, not bits), 0 on the
    // host (the reference's bits), -1 on the GPU for 16-bit storage, else on the host
    int fit_on_device = -1;
    // Batched form: the real pass and the null runs go through the stacking kernel in
    // groups of passes_per_launch passes per launch (0: the searcher's tuned value, else
    // every pass in one group, within the memory budget; 1: one pass at a time with a
    // plan per pass). segments: velocity-block segments per tile (0: the planner's
    // choice). Neither changes any result bit.
    int passes_per_launch = 1;
    int segments = 0;
    // Stacking work shape: output rows per tile (8, 16 or 32; 0 = tuned, else
    // automatic) and the launch form (1: one persistent launch per group whose
    // workgroups take work items from device counters; 0: one workgroup per work item,
    // 16-row tiles only; -1 = tuned, else automatic, see resolve_shape).
    int tile_rows = 0;
    int queue = -1;
    // Stage builds of a group in parallel on this many streams (1-4; 0 = by frame size:
    // 4 up to 16 M samples per pass, 2 up to 64 M, else 1).
    int build_lanes = 0;
    // Frame storage of the stacking passes: 0 fp32 (the exact form), 1 fp16, 2 bf16
    // (a power-of-two scale per frame, fp32 accumulation).
    int storage = 0;
    // Null-peak screen (gauss null model): the strongest null_tested peaks of each null
    // run go through the artefact tests at their grid position and motion and artefacts
    // stay out of the pooled peaks (0 = test none); max_frame_fraction is the
    // single-frame test's limit. reg_dx / reg_dy: the frames' registration offsets
    // (detector = registered + offset), one per frame, or empty (no detector test).
    int null_tested = 3000;
    double max_frame_fraction = 0.25;
    std::vector<double> reg_dx, reg_dy;
    // Candidate stage: 1 = the selected peaks are refined into candidates (refine and
    // finish_candidates, src/search.cpp:246-395, 627-675), 0 = the run stops at the
    // selection; half_tolerance is the halves test's allowed shortfall (sigma).
    int candidates = 1;
    double half_tolerance = 3.0;
    // Near-Earth search (SearchParams::neo_min_dist_au > 0): curved tracks x + v t + q
    // par(t) over the inverse distances q of inverse_distance_grid (src/parallax.cpp:
    // 167-185) times the velocity grid, motion j = k nv + iv for q = Q[k], v = grid[iv]
    // (search.cpp:727-733). par_x / par_y: each frame's parallax displacement at 1 AU
    // in pixels, less its least-squares line in t (Prepared::par_x / par_y,
    // set_parallax_model, src/parallax.cpp:132-156), one per frame; needed when
    // neo_min_dist_au > 0.
    double neo_min_dist_au = 0.0;
    int neo_max_trials = 64;
    std::vector<double> par_x, par_y;
    // whether a run has a candidate stage (candidates on and room for at least one)
    bool refines() const { return candidates != 0 && max_candidates > 0; }
};

struct PeakList {
    std::vector<int> x, y, vidx;
    std::vector<float> snr;
};

// A candidate (Candidate, include/fainttrail/search.hpp:131-154): the fields a search
// without a WCS fills (0-based position at the reference epoch).
struct SearchCandidate {
    int id = 0;
    double x = 0.0, y = 0.0, vx_pxh = 0.0, vy_pxh = 0.0;
    double pos_err_px = 0.0, vel_err_pxh = 0.0;
    double snr = 0.0, snr_coarse = 0.0, snr_half1 = 0.0, snr_half2 = 0.0;
    double max_frame_fraction = 0.0, false_alarms_per_run = 0.0;
    double flux = 0.0, mag_inst = 0.0;
    double inv_dist_au = 0.0, inv_dist_err_au = 0.0;
    bool ok = false;
    std::string flags;
};

struct SearchOutput {
    double threshold = 0.0;
    int method = 0;  // 0 null-fit gauss, 1 null-fit linear, 2 null-quantile, 3 fixed, 4 gaussian
    double null_a = 0.0, null_k = 0.0, null_inv_s2 = 0.0, null_b = 0.0;
    double dv = 0.0, t_ref_mjd = 0.0, span_h = 0.0;
    std::vector<double> t_h;
    std::vector<std::pair<double, double>> grid;
    std::vector<double> inv_dist{0.0};  // the inverse distances (1/AU) of the motions; {0}: straight tracks
    double dq = 0.0;                    // their step (0 for straight tracks)
    // motions of a pass: every (inverse distance, velocity) pair
    size_t motions() const { return grid.size() * inv_dist.size(); }
    std::vector<float> pooled;  // null peak S/N values, run after run (artefacts left out)
    int screened = 0, artefacts = 0;  // null peaks through the artefact tests, and rejected
    PeakList peaks;             // real-pass peaks, sorted, at most 100000
    PeakList selected;          // peaks at or above the threshold, at most max_candidates
    std::vector<SearchCandidate> candidates;  // sorted as the reference sorts them, ids from 1
    std::vector<std::pair<std::string, double>> timings;
    int passes_per_launch = 1, segments = 0;  // what the run used
    int tile_rows = 16, queue = 0;
};

// Device buffers of one group of stacking passes (the batched form).
struct PassSet {
    int cap = 0;                   // passes
    long long plane = 0, frames = 0, frame_floats = 0, tab_ints = 0, pk_cap = 0;
    long long layout = -1;         // the frame layout the border zeros were written for
    float* filt = nullptr;         // cap x frames x frame_floats (a padded frame each)
    unsigned long long* keys = nullptr;
    float* maxm = nullptr;         // cap x plane
    int* best = nullptr;
    int* tabs = nullptr;           // cap x tab_ints: shift, box, lds_off rows of the pass
    int* perm = nullptr;           // cap x frames
    int* counts = nullptr;         // cap peak counts
    int* pk = nullptr;             // cap x 4 pk_cap: x, y, snr bits, vidx
    float* scl = nullptr;          // cap x frames: scales of 16-bit frames
    unsigned* maxb = nullptr;      // cap x frames: scratch of the 16-bit encoder
    int* h_counts = nullptr;       // pinned
    int* queue_ws = nullptr;       // the persistent launch's 8 counters
    void release();
    void reserve(int cap, long long frames, long long frame_floats, long long plane, long long tab_ints,
                 long long pk_cap);
    // device bytes the set's buffers hold (the sizes asked of hipMalloc)
    double device_bytes() const;
};

class DeviceSearch {
public:
    DeviceSearch(int n, int w, int h);
    ~DeviceSearch();
    DeviceSearch(const DeviceSearch&) = delete;
    DeviceSearch& operator=(const DeviceSearch&) = delete;

    // raw: n frames (device, stride floats apart) sorted by time; mjd: their mid-exposure
    // MJDs (host). d_max_snr / d_best_v (optional) receive the real pass's maps.
    SearchOutput run(const float* raw, long long stride, const double* mjd, const SearchConfig& cfg, float* d_max_snr,
                     int* d_best_v, hipStream_t s);

    int n, w, h;
    int tuned_passes = 0, tuned_segments = 0, tuned_tile_rows = 0;  // from tune(); 0 = none
    int tuned_queue = -1;                                           // from tune(); -1 = none

    // The threshold of a run's pooled null peaks fitted again with the FP64 tail-fit
    // sums on the host (on_device false) or the GPU; method receives the method index.
    double refit(const SearchConfig& cfg, const SearchOutput& last, bool on_device, int* method, hipStream_t s);

    // Device bytes the searcher holds now (every buffer it allocated and keeps, as
    // asked of hipMalloc; buffers only grow between runs).
    double device_bytes() const;

    // Persistent stacking launches of the batched form go through this fuser (the
    // search server's), which may stack them together with other jobs' (null: launched
    // directly).
    void set_fuser(class StackFuser* f) { fuser_ = f; }

    // Times the batched run over candidate group sizes and segment counts, then over
    // the launch forms and tile heights, on these frames and keeps the fastest
    // (tuned_passes, tuned_segments, tuned_tile_rows, tuned_queue). Returns the seconds
    // per search of the choice.
    double tune(const float* raw, long long stride, const double* mjd, const SearchConfig& cfg, int reps,
                hipStream_t s);

private:
    SearchOutput run_sequential(const float* raw, long long stride, const SearchConfig& cfg, SearchOutput out,
                                float* d_max_snr, int* d_best_v, hipStream_t s);
    SearchOutput run_batched(const float* raw, long long stride, const SearchConfig& cfg, SearchOutput out,
                             float* d_max_snr, int* d_best_v, hipStream_t s);
    void finish(const SearchConfig& cfg, int runs, int motions, SearchOutput& out, hipStream_t s);
    void fit_threshold(const SearchConfig& cfg, int runs, int motions, bool on_device, SearchOutput& out, hipStream_t s);
    // The candidate stage: the selected peaks refined on the device against the real
    // pass's frames, residuals and mask, then finished on the host.
    void refine(const SearchConfig& cfg, SearchOutput& out, hipStream_t s);
    // The real pass's buffers of the candidate stage (allocated on first use).
    void reserve_real();
    // The null-peak screen of one or more passes: the artefact flags of the first
    // min(count, null_tested) peaks of each pass, tested on that pass's frames.
    struct ScreenPass {
        const void* frames;      // frame i element (x, y) at frames + i fstride + y pitch + x
        long long fstride;
        int pitch;
        int storage;
        const float* scale;      // 16-bit storage: per-frame scales (device)
        int pass;                // index of the pass's frame times in the run's table
        const PeakList* peaks;
    };
    std::vector<std::vector<unsigned char>> screen_passes(const SearchConfig& cfg, const std::vector<ScreenPass>& passes,
                                                          hipStream_t s, SearchOutput& out);
    // The run's screen tables (velocities, sqrt(c), detector steps, registration
    // offsets, frame times of every pass), uploaded before the first screen of a run.
    void screen_tables(const SearchConfig& cfg, const SearchOutput& out, const std::vector<std::vector<double>>& times,
                       hipStream_t s);
    PeakList peaks(const float* m, const int* b, float floor, int radius, int edge, size_t max_peaks, hipStream_t s);
    DeviceStage stage_;
    float* sub_ = nullptr;
    float* filt_ = nullptr;
    float* max_ = nullptr;
    int* best_ = nullptr;
    float* nmax_ = nullptr;  // null-run maps
    int* nbest_ = nullptr;
    int* pk_ = nullptr;  // count + 4 arrays
    size_t pk_cap_ = 0;
    float* sky_ = nullptr;  // the run's static sky (the screen's blank test)
    // candidate stage: the real pass's filtered frames (fp32), residuals and mask; the
    // refinement's tables and figures (device + pinned staging)
    float* real_ = nullptr;
    float* res_ = nullptr;
    float* rmask_ = nullptr;
    char* ref_dev_ = nullptr;
    char* ref_host_ = nullptr;
    size_t ref_cap_ = 0;
    // null screen: run tables (device + pinned staging), peaks and flags
    double* scr_tab_ = nullptr;
    double* scr_tab_host_ = nullptr;
    size_t scr_tab_cap_ = 0;
    char* scr_pk_ = nullptr;       // frames descriptors, then x, y, velocity index, pass of each peak
    char* scr_pk_host_ = nullptr;
    unsigned char* scr_art_ = nullptr;
    unsigned char* scr_art_host_ = nullptr;
    size_t scr_cap_ = 0, scr_art_cap_ = 0;
    int scr_nv_ = 0, scr_nsteps_ = 0, scr_edge_r_ = 0;
    int scr_nq_ = 1, scr_passes_ = 0;  // inverse distances; passes whose times (and parallax) the tables hold
    bool scr_reg_ = false;
    bool scr_curved_ = false;          // near-Earth: per-pass parallax and the inverse distances in the tables
    PassSet sets_[2];
    hipStream_t build_stream_ = nullptr, copy_stream_ = nullptr;
    // more build lanes for small frames (each its own stage scratch, build target and
    // stream), created on first use: several passes of a group build at the same time
    struct BuildLane {
        std::unique_ptr<DeviceStage> stage;
        float* filt = nullptr;
        hipStream_t stream = nullptr;
        hipEvent_t built[2] = {nullptr, nullptr};
    };
    std::vector<BuildLane> lanes_;  // lanes 1, 2, ... (lane 0: stage_, filt_, build_stream_)
    hipEvent_t ev_built_[2] = {nullptr, nullptr}, ev_stacked_[2] = {nullptr, nullptr};
    hipEvent_t ev_counts_[2] = {nullptr, nullptr};
    hipStream_t ring_stream_ = nullptr;  // the ring tiles' launches, beside the interior ones
    hipEvent_t ev_ring_in_[2] = {nullptr, nullptr}, ev_ring_out_[2] = {nullptr, nullptr};
    hipEvent_t ev_start_ = nullptr;
    // per-run tables of the batched form (plan rows, frame orders, time ranks, stacking
    // descriptors): packed in pinned host memory and sent in one copy to a device buffer,
    // both grown on demand and reused; ev_tables_ marks the previous run's last use
    void* tab_host_ = nullptr;
    void* tab_dev_ = nullptr;
    size_t tab_bytes_ = 0;
    hipEvent_t ev_tables_ = nullptr, ev_upload_ = nullptr;
    bool tables_used_ = false;
    // the batched form's host plan of the last run, reused while its inputs (frame times,
    // velocity grid, coverage count) are the same: planning is the run's largest host cost
    // at small frames
    std::unique_ptr<StackPlan> plan_cache_;
    std::vector<int32_t> plan_key_dx_, plan_key_dy_;
    int plan_key_mc_ = -1;
    // pinned host staging of a group's peak lists (x, y, snr, vidx of each pass)
    int* pk_host_ = nullptr;
    size_t pk_host_cap_ = 0;
    class StackFuser* fuser_ = nullptr;
};

// Build lanes of the batched form for n frames of `plane` pixels and `passes` passes
// (asked > 0: that many, at most 4 and at most the passes).
int build_lanes(int n, long long plane, int passes, int asked = 0);

// Host pieces of the loop, restated from the reference.
std::vector<std::pair<double, double>> velocity_grid(double span_h, double vmin, double vmax, double tol, double* dv);
// inverse_distance_grid (src/parallax.cpp:167-185) for cfg's near-Earth settings and
// parallax displacements: {0} when the search is off or the curvature is below the
// grid tolerance; dq (may be null) receives the step (0 for {0}).
std::vector<double> inverse_distance_grid(const SearchConfig& cfg, int n, double* dq);
std::vector<std::vector<double>> null_time_orders(const std::vector<double>& t_h, int runs, uint64_t seed);
// The same shuffles as index permutations: null_time_orders(t, runs, seed)[k][i] ==
// t[null_time_perms(n, runs, seed)[k][i]].
std::vector<std::vector<int>> null_time_perms(int n, int runs, uint64_t seed);
// find_peaks' host order (search.cpp:105-135) of c device peaks: row-major, then a
// std::sort by descending S/N, then truncation to max_peaks.
PeakList order_peaks(int c, const int* x, const int* y, const float* snr, const int* vidx, size_t max_peaks);
// pool_null_peaks (search.cpp:690-693): the run's peaks in order, artefacts left out.
void pool_null(const PeakList& pk, const std::vector<unsigned char>& art, SearchOutput& out);
void set_threshold(const SearchConfig& cfg, int width, int height, double fwhm, size_t n_motions, int runs,
                   const std::vector<float>& pooled, SearchOutput& out, const double* device_sums);
// expected_null_peaks (search.cpp:696-700): noise peaks per run above snr under the
// fitted null tail (0 without a fit).
double expected_null_peaks(const SearchOutput& out, double snr);
// refine's host tail (search.cpp:319-391) and finish_candidates (:627-675) for the
// selected peaks of out: fig holds kRefOut device figures per selected peak, contrib n
// per-frame samples each. Returns the candidates in the reference's order, ids from 1.
std::vector<SearchCandidate> finish_candidates(const SearchConfig& cfg, const SearchOutput& out, const double* fig,
                                               const float* contrib);

}  // namespace ftk
