/* ftk_hip.h - C-ABI of the MI355X (gfx950) FaintTrail search backend.
 *
 * Plain C: POD structs, plain pointers and sizes, integer status codes and a
 * caller-supplied error buffer. No C++ or torch types cross this boundary.
 *
 * Status codes mirror the reference C API (include/fainttrail/fainttrail.h:26-28):
 * FTK_OK, FTK_ERROR (err receives the message; validation messages follow
 * validate_pass, src/backend.cpp:89-108) and FTK_CANCELLED (the progress callback
 * returned 0; the reference throws Cancelled, include/fainttrail/types.hpp:70-72).
 *
 * Layouts:
 *   frames  n frames of w x h row-major fp32 (the reference's Image, types.hpp:20-34)
 *   shifts  motion-major int32: motion j, frame i at j * n + i (ShiftTable,
 *           include/fainttrail/backend.hpp:56-69)
 *   results max_snr fp32 and best_v int32, w x h row-major (StackResult,
 *           backend.hpp:99-105): 0 / -1 where fewer than min_count frames cover.
 */
#ifndef FTK_HIP_H
#define FTK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTK_OK 0
#define FTK_ERROR 1
#define FTK_CANCELLED 2
#define FTK_TIMEOUT 3 /* ftk_server_wait: the job has not ended yet */

/* Progress in [0,1] and a stage name; return 0 to cancel.
 * Same shape as ft_progress_fn (fainttrail.h:71) and ProgressFn (types.hpp:76). */
typedef int (*ftk_progress_fn)(double fraction, const char* stage, void* user);

/* Backend options, "hip:device=N,max_mb=M,launch_adds=A,pad=P,tile_rows=R,queue=Q".
 * Replaces CudaOptions / parse_cuda_options (src/cuda_plan.hpp:27-32,
 * src/cuda_plan.cpp:14-33): device -1 = FAINTTRAIL_HIP_DEVICE or 0; max_mb caps the
 * device memory for frame rows (0 = 80 % of free memory; small values force row
 * bands); launch_adds = stacked units per kernel launch (progress and cancel
 * granularity; default 4e12, about 0.25 s on MI355X); pad = 1 (default) keeps the
 * frames in HBM inside a zero border that holds every shifted read (when that costs
 * at most twice their size), 0 = plain rows. tile_rows and queue have no reference
 * counterpart: tile_rows = output rows per stacking tile (8, 16 or 32; 0 = automatic,
 * 16; 64 columns wide); queue = 1 stacks each launch as one persistent kernel whose
 * workgroups take (kind, tile, velocity segment) work items from device counters, 0 =
 * one workgroup per work item in up to three launches (16-row tiles only), -1
 * (default) = automatic: persistent for 16-bit frames and for fp32 frames of at least
 * 1024 tiles of 64 x 16. ftk_plan_tune picks both by timing. The results do not
 * depend on pad, tile_rows or queue. */
typedef struct ftk_options {
    int device;
    double max_mb;
    double launch_adds;
    int pad;
    int tile_rows;
    int queue;
} ftk_options;

/* Parses the part after "hip:" (NULL or "" = defaults). Unknown keys or
 * non-numbers are errors, as in parse_cuda_options. */
int ftk_parse_options(const char* spec, ftk_options* out, char* err, size_t err_len);

/* Number of usable HIP devices (0 when none; why receives the reason).
 * Replaces cuda_device_count (src/backend_cuda.cu:198-209). */
int ftk_device_count(char* why, size_t why_len);

/* Human-readable device description for SearchBackend::description()
 * (backend_cuda.cu:89-94). */
int ftk_device_description(int device, char* buf, size_t len);

/* One stacking pass: the inputs of StackPass (backend.hpp:71-78) with the shifts
 * always given as a table (a straight-line grid arrives as linear_shift_table,
 * backend.cpp:114-123, which the contract defines as identical). */
typedef struct ftk_pass {
    int n, w, h;
    const float* const* frames; /* n host pointers, each w * h fp32, read-only */
    int nv;                     /* trial motions                               */
    const int32_t* dx;          /* nv * n, motion-major                        */
    const int32_t* dy;
    int min_count;
} ftk_pass;

/* Sub-timings reported in StackResult::timings (backend_cuda.cu:189-192). */
typedef struct ftk_timings {
    double upload_s, kernel_s, download_s, setup_s;
} ftk_timings;

/* SearchBackend::run for the "hip" backend (backend.hpp:107-117; replaces
 * CudaBackend::run, backend_cuda.cu:99-195). Synchronous; host in, host out.
 * max_snr and best_v are caller-allocated w * h arrays. Bit-identical to the CPU
 * backend (backend_cpu.cpp:77-204). timings / progress may be NULL. */
int ftk_stack_pass(const ftk_pass* pass, const ftk_options* opt, float* max_snr, int32_t* best_v,
                   ftk_timings* timings, ftk_progress_fn progress, void* user, char* err, size_t err_len);

/* ------------------------------------------------------------------ device-resident form
 * No reference counterpart: the reference uploads the frames on every pass
 * (backend_cuda.cu:146-150). A plan holds the velocity blocks, shift tables and
 * scratch on the device; ftk_plan_run stacks frames that are already in HBM.
 * Used by the batched search and by callers that keep frames resident. */
typedef struct ftk_plan ftk_plan;

int ftk_plan_create(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count,
                    const ftk_options* opt, ftk_plan** out, char* err, size_t err_len);
/* d_frames: n frames of w x h fp32 on the device, frame i at d_frames + i * frame_stride.
 * d_max_snr / d_best_v: w x h device outputs. stream: a hipStream_t (NULL = default).
 * Asynchronous with respect to the host. The edge tiles run on a stream the plan owns,
 * ordered after the work already on `stream` and before what follows on it, so the
 * call behaves as stream-ordered work; runs of one plan must not overlap (they share
 * its scratch). */
int ftk_plan_run(ftk_plan* plan, const float* d_frames, size_t frame_stride, float* d_max_snr, int32_t* d_best_v,
                 void* stream, char* err, size_t err_len);
/* Units (motions x frames x pixels) one ftk_plan_run stacks, and the number of
 * kernel launches it makes. */
double ftk_plan_units(const ftk_plan* plan);
int ftk_plan_blocks(const ftk_plan* plan);
void ftk_plan_destroy(ftk_plan* plan);

/* ------------------------------------------------------------------ precision form
 * No reference counterpart (the reference stacks fp32 frames only). Frames kept in HBM
 * as fp16 or bf16 values q with a power-of-two scale s_i per frame, s_i = 2^(e - 14)
 * for max|g_i| = m 2^e (m in [0.5, 1)), q = round-to-nearest-even(g / s_i); the
 * stacking kernel widens q * s_i exactly and accumulates in fp32 in the contract's
 * order, so a stored pass equals the fp32 pass over the decoded frames bit for bit. */
#define FTK_STORE_F32 0
#define FTK_STORE_F16 1
#define FTK_STORE_BF16 2
/* Encodes n fp32 frames (device, frame_stride floats apart) into the plan's 16-bit
 * storage (the fp32 frames may be freed afterwards). Asynchronous. */
int ftk_plan_store(ftk_plan* plan, const float* d_frames, size_t frame_stride, int storage, void* stream, char* err,
                   size_t err_len);
/* A stacking pass over the stored 16-bit frames. Asynchronous. */
int ftk_plan_run_stored(ftk_plan* plan, float* d_max_snr, int32_t* d_best_v, void* stream, char* err, size_t err_len);
/* The n per-frame scales of the stored frames (host copy; synchronises the device). */
int ftk_plan_scales(const ftk_plan* plan, float* scales);

/* Autotuner of a plan's work shape, no reference counterpart: times the plan's pass
 * (reps runs after one warm-up) on these fp32 frames (as ftk_plan_run), or on the
 * stored 16-bit frames when d_frames is NULL (as ftk_plan_run_stored), for each launch
 * form and tile height the kernels take (one workgroup per work item at 16 rows; the
 * persistent launch at 8, 16 and 32 rows), and keeps the fastest for that storage's
 * later runs. Writes the choice and its seconds per pass (outputs may be NULL); the
 * maps receive the pass's result (the same for every shape). Synchronous. */
int ftk_plan_tune(ftk_plan* plan, const float* d_frames, size_t frame_stride, float* d_max_snr, int32_t* d_best_v,
                  int reps, void* stream, int* tile_rows, int* queue, double* seconds, char* err, size_t err_len);
/* The stacking contract evaluated in FP64 on the GPU, the precision reference for
 * stored and fp32 passes: output rows [y0, y1) of n fp32 frames (plain rows of w),
 * sums of the covering frames in double, times 1/sqrt(c) in double, strict maximum
 * over the motions (ties to the lowest index). dx, dy: host tables, motion-major.
 * d_max_snr (double) / d_best_v: (y1 - y0) x w device outputs. Synchronous. */
int ftk_stack_fp64(int n, int w, int h, const float* d_frames, size_t frame_stride, int nv, const int32_t* dx,
                   const int32_t* dy, int min_count, int y0, int y1, double* d_max_snr, int32_t* d_best_v, void* stream,
                   char* err, size_t err_len);
/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * ftk_stack_fp64 with the same kernels and results, but it allocates nothing. The shift
 * tables d_dx / d_dy are device arrays (nv * n, motion-major) and the per-motion-chunk
 * partial maxima live in a fixed 64 MiB device array of the library (one per GPU, held
 * until the process ends) instead of memory allocated per call; rows go in strips that
 * fit it. Fails (FTK_ERROR, nothing launched) when one row of partials,
 * w * ceil(nv / 128) * 12 bytes, exceeds 64 MiB. Asynchronous on stream. Every call on
 * one GPU shares that array: calls that overlap on one GPU (different streams or
 * threads) race; calls in one stream are ordered. */
int ftk_stack_fp64_noalloc(int n, int w, int h, const float* d_frames, size_t frame_stride, int nv,
                           const int32_t* d_dx, const int32_t* d_dy, int min_count, int y0, int y1, double* d_max_snr,
                           int32_t* d_best_v, void* stream, char* err, size_t err_len);

/* ------------------------------------------------------------------ preprocessing stage
 * The time-order-dependent preprocessing of a search, build_stage
 * (src/preprocess.cpp:286-486), and the per-frame background of prepare_core
 * (:568-581), on the device and bit-identical to the reference. It replaces the host
 * prepare() / with_times() calls (internal.hpp:55-58, preprocess.cpp:488-521) that
 * feed every stacking pass: the null runs rebuild this stage with permuted times. */
typedef struct ftk_stage ftk_stage;

/* SearchParams fields that build_stage reads (include/fainttrail/search.hpp:18-90);
 * fwhm_px is the PSF FWHM already estimated on the real data. */
typedef struct ftk_stage_params {
    double fwhm_px;
    double clip_sigma;      /* 8   */
    double cr_sigma;        /* 5   */
    int static_exclude;     /* -1 = max(1, lround(n / 8)) */
    double mask_scatter;    /* 2.5 */
    double weight_scatter;  /* 1.5 */
} ftk_stage_params;

/* Device scratch for n frames of w x h (at most 255 frames). */
int ftk_stage_create(int n, int w, int h, int device, ftk_stage** out, char* err, size_t err_len);
void ftk_stage_destroy(ftk_stage* stage);
/* d_raw (n frames, stride raw_stride floats) -> d_sub: float(v - median background),
 * NaN-grown blank areas; background (host, n doubles, may be NULL). */
int ftk_stage_background(ftk_stage* stage, const float* d_raw, size_t raw_stride, float* d_sub, size_t sub_stride,
                         double* background, void* stream, char* err, size_t err_len);
/* build_stage on the device for the frame times t_h (host, n; only their order
 * matters): d_filtered receives the normalised matched-filtered frames (the frames of
 * the stacking pass). Optional outputs (NULL to skip): d_static_sky (w x h device),
 * sigma (n host doubles), d_mask (w x h device), cr_pixels (n host ints). When sigma
 * or cr_pixels is given the call synchronises the stream. */
int ftk_stage_build(ftk_stage* stage, const float* d_sub, size_t sub_stride, const double* t_h,
                    const ftk_stage_params* params, float* d_filtered, size_t filt_stride, float* d_static_sky,
                    double* sigma, float* d_mask, int32_t* cr_pixels, void* stream, char* err, size_t err_len);

/* ------------------------------------------------------------------ device-resident search
 * The reference's blind search over prepared frames, search() (src/search.cpp:716-809):
 * prepare (background + build_stage), the real stacking pass, find_peaks, null_runs
 * time-shuffled null runs (with_times rebuild + pass + peaks, search.cpp:761-782), the
 * null-tail threshold (set_threshold, :573-612) and select_peaks (:614-620), with the
 * frames resident on the device from the first step to the last, the strongest peaks
 * of each null run screened for artefacts on the device (is_artefact_peak,
 * search.cpp:229-243, 773-778), and the candidate stage: the selected peaks refined on
 * the device (refine, search.cpp:246-395: local search, sub-step parabolas, halves,
 * mask, PSF photometry on the stacked residuals, edge and detector tests) and finished
 * on the host (finish_candidates, :627-675: duplicates, ghosts, flags, false-alarm
 * rates, order, ids). Straight tracks only (no near-Earth distance grid); no WCS, so
 * no sky coordinates. */
typedef struct ftk_search_params {
    double vmin_pxh, vmax_pxh, grid_tol_px;
    double fwhm_px; /* PSF FWHM, must be > 0 (estimated once, outside this path) */
    double clip_sigma, cr_sigma;
    int static_exclude;
    double mask_scatter, weight_scatter, min_coverage;
    double threshold_snr; /* 0 = calibrate with the null runs */
    double min_snr_floor;
    int null_runs;
    double target_false_per_run;
    int max_candidates, edge_px;
    unsigned long long seed;
    /* Where the FP64 null-tail fit sums run: 1 on the GPU (same method; the threshold
     * agrees with the host fit within the host sums' rounding, ~m eps relative for m
     * fitted peaks), 0 on the host (the reference's bits), -1 (default) on the GPU for
     * 16-bit storage, else on the host. */
    int fit_on_device;
    /* Batched form (no result bit depends on these): stacking passes per launch, the
     * real pass and the null runs in groups of this many (1: one pass at a time, a
     * plan per pass; 0: the searcher's tuned value, else every pass in one group, within
     * half the free device memory), and velocity-block segments per tile (0: the
     * planner's choice, or the tuned value). */
    int passes_per_launch;
    int segments;
    /* Frame storage of the stacking passes: FTK_STORE_F32 (the exact form), or
     * FTK_STORE_F16 / FTK_STORE_BF16 (see the precision form below). */
    int storage;
    /* SearchParams::null_tested (default 3000; 0 = test none): the strongest null
     * peaks of each null run that go through the artefact tests (single frame, field
     * edge, detector-fixed) at their grid position and motion; artefacts stay out of
     * the null tail fit. With 16-bit storage the tests read the decoded frames. */
    int null_tested;
    double max_frame_fraction; /* the single-frame test's limit (default 0.25) */
    /* Registration offsets of the frames (n host doubles each, detector = registered +
     * offset; Frame::reg_dx / reg_dy), or NULL: with them the detector-fixed test
     * runs (search.cpp:196-225), as in the reference when every frame carries one.
     * Copied by the call. */
    const double* reg_dx;
    const double* reg_dy;
    /* Candidate stage: 1 (default) = refine the selected peaks into candidates, 0 =
     * stop at the selection. half_tolerance: SearchParams::half_tolerance (default 3),
     * the halves test's allowed shortfall in sigma. */
    int candidates;
    double half_tolerance;
    /* Stacking work shape (no result bit depends on these): output rows per tile (8,
     * 16 or 32; 0 = the searcher's tuned value, else automatic) and the launch form
     * (1: one persistent launch per group of passes whose workgroups take work items
     * from device counters; 0: one workgroup per work item, 16-row tiles only; -1,
     * default: tuned, else automatic as the plan options' queue -1). */
    int tile_rows;
    int queue;
    /* Stage builds of a group of passes run in parallel on this many streams, each fed
     * by a host thread of its own (1-4; 0 = by frame size). No result bit depends on it. */
    int build_lanes;
    /* Near-Earth search (SearchParams::neo_min_dist_au / neo_max_trials, search.hpp:
     * 86-88; default 0 = off, 64): with neo_min_dist_au > 0 the search runs curved
     * tracks x + v t + q par(t) over the inverse distances q of inverse_distance_grid
     * (src/parallax.cpp:167-185) times the velocity grid, motion j = k nv + iv
     * (search.cpp:727-733), and candidates carry inv_dist_au / inv_dist_err_au. par_x /
     * par_y: each frame's parallax displacement at 1 AU in pixels, less its
     * least-squares line in t (Prepared::par_x / par_y as set_parallax_model leaves
     * them, src/parallax.cpp:132-156), n host doubles each, or NULL when off; copied by
     * the call. */
    double neo_min_dist_au;
    int neo_max_trials;
    const double* par_x;
    const double* par_y;
} ftk_search_params;

typedef struct ftk_search_summary {
    double threshold_snr;
    int method; /* 0 null-fit gauss, 1 null-fit linear, 2 null-quantile, 3 fixed, 4 gaussian */
    double null_a, null_k, null_inv_s2, null_b;
    double dv_pxh, t_ref_mjd, span_h;
    int n_velocities, n_peaks, n_selected, n_pooled;
    double t_prepare, t_stack_real, t_null_rebuild, t_stack_null, t_peaks, t_threshold; /* seconds */
    /* batched form: host planning (with the background step) and the pipelined passes
     * (stage builds, stacking and peaks of every pass); the group size and segment
     * count the run used */
    double t_plan, t_passes;
    int passes_per_launch, segments;
    int n_screened, n_artefacts; /* null peaks through the artefact tests; of those, left out */
    int n_candidates;            /* candidates after the duplicate test */
    double t_refine;             /* seconds of the candidate stage */
    int tile_rows, queue;        /* the stacking work shape the run used */
    int n_distances;             /* inverse distances of the motions (1: straight tracks) */
    double dq_inv_au;            /* their step (0: straight tracks) */
} ftk_search_summary;

/* A candidate: ft_candidate (include/fainttrail/fainttrail.h:56-69) with the position
 * 0-based as in Candidate (include/fainttrail/search.hpp:131-154; ft_get_candidate,
 * src/c_api.cpp:218-246, adds 1 for FITS), and Candidate's inverse distance fields
 * (0 on straight tracks). ra_deg .. pa_deg stay 0: this path has no WCS. */
typedef struct ftk_candidate {
    int id;
    int ok; /* 1 = passed every consistency test */
    double x, y;
    double vx_pxh, vy_pxh;
    double pos_err_px, vel_err_pxh;
    double snr, snr_coarse, snr_half1, snr_half2;
    double max_frame_fraction;
    double false_alarms_per_run;
    double flux, mag_inst;
    double ra_deg, dec_deg;
    double rate_arcsec_min, pa_deg;
    double inv_dist_au, inv_dist_err_au;
    char flags[64]; /* comma separated, as Candidate::flags */
} ftk_candidate;

typedef struct ftk_searcher ftk_searcher;

/* SearchParams defaults (include/fainttrail/search.hpp:18-90), fwhm_px = 3. */
void ftk_search_params_default(ftk_search_params* params);
int ftk_searcher_create(int n, int w, int h, int device, ftk_searcher** out, char* err, size_t err_len);
void ftk_searcher_destroy(ftk_searcher* searcher);
/* d_raw: n registered frames on the device, sorted by time (as ft_search sorts them,
 * c_api.cpp:178-180), stride floats apart; mjd: their mid-exposure MJDs (host).
 * d_max_snr / d_best_v (optional, w x h device) receive the real pass's maps. */
int ftk_searcher_run(ftk_searcher* searcher, const float* d_raw, size_t stride, const double* mjd,
                     const ftk_search_params* params, ftk_search_summary* summary, float* d_max_snr,
                     int32_t* d_best_v, void* stream, char* err, size_t err_len);
/* Autotuner of the batched form: runs the search on these frames (params as for
 * ftk_searcher_run) over candidate group sizes and velocity segments per tile, then
 * over the launch forms and tile heights (one workgroup per work item at 16 rows;
 * persistent at 8, 16, 32 rows) at the best of those, reps times each, and keeps the
 * fastest for later runs with passes_per_launch = 0 / segments = 0 / tile_rows = 0 /
 * queue = -1 (a tile height or launch form the params fix is kept). Writes the choice
 * and its seconds per search (any output may be NULL). No extension of the
 * reference: its searches run one pass at a time. */
int ftk_searcher_tune(ftk_searcher* searcher, const float* d_raw, size_t stride, const double* mjd,
                      const ftk_search_params* params, int reps, int* passes_per_launch, int* segments,
                      int* tile_rows, int* queue, double* seconds, void* stream, char* err, size_t err_len);
/* Peaks of the last run: which 0 = real-pass peaks (sorted as find_peaks sorts them),
 * 1 = selected peaks, 2 = pooled null peak S/N values (snr only). Copies up to cap
 * entries into the non-NULL arrays and returns the total count. */
int ftk_searcher_peaks(const ftk_searcher* searcher, int which, int cap, int32_t* x, int32_t* y, float* snr,
                       int32_t* vidx);
/* The null-tail threshold of the last run's pooled null peaks fitted again (params as
 * for that run) with the FP64 fit sums on the host (on_device 0) or the GPU (1):
 * threshold and method index as in ftk_search_summary. */
int ftk_searcher_refit(ftk_searcher* searcher, const ftk_search_params* params, int on_device, double* threshold,
                       int* method, void* stream, char* err, size_t err_len);
/* Candidates of the last run, in the reference's order (ok ones first, then by S/N;
 * ft_candidate_count / ft_get_candidate, src/c_api.cpp:213-246): writes up to cap
 * records and returns the count, -1 for a null searcher. */
int ftk_searcher_candidates(const ftk_searcher* searcher, int cap, ftk_candidate* out);
/* find_peaks (src/search.cpp:105-135) over device maps (w x h max_snr and best_v):
 * pixels in [edge, w - edge) x [edge, h - edge) at or above floor that are the maximum
 * of their (2 radius + 1)^2 window (strict against later pixels, row-major, and
 * non-strict against earlier ones: one peak per plateau), sorted by descending S/N
 * after a row-major scan (std::sort, as the reference) and truncated to max_peaks.
 * Writes up to cap peaks into the non-NULL arrays and the total into *count.
 * Synchronous on stream. */
int ftk_find_peaks(const float* d_max_snr, const int32_t* d_best_v, int w, int h, float floor, int radius, int edge,
                   int max_peaks, int cap, int32_t* x, int32_t* y, float* snr, int32_t* vidx, int* count,
                   void* stream, char* err, size_t err_len);

/* ------------------------------------------------------------------ serving form
 * No reference counterpart (one search per process there). A device-resident request
 * loop: submitted searches queue in arrival order; a free one of `slots` worker slots
 * admits the first queued job whose HBM reservation fits the budget beside the running
 * jobs and the idle searchers kept for reuse (evicted, oldest first, when room is
 * needed). A job may overtake the head of the queue only while the head has been
 * overtaken fewer than 16 times. The reservation is the job's footprint estimate (its
 * searcher and batched pass buffers), or the device bytes a reused searcher of its
 * shape already holds if that is more; an ended job's searcher is kept at the bytes
 * it holds. Each slot runs its job's batched search on its own streams, so jobs
 * overlap on the device. A job's results depend on its inputs only. */
typedef struct ftk_server ftk_server;
/* budget_mb <= 0: 85 % of the free device memory. */
int ftk_server_create(int device, double budget_mb, int slots, ftk_server** out, char* err, size_t err_len);
/* Finishes the running jobs; queued jobs fail. */
void ftk_server_destroy(ftk_server* server);
/* d_raw: n frames (device, stride floats apart) sorted by time, kept valid and
 * unchanged until the job ends; mjd: n times (copied). params as for
 * ftk_searcher_run (passes_per_launch 0: up to 8, within the budget). */
int ftk_server_submit(ftk_server* server, const float* d_raw, size_t stride, int n, int w, int h, const double* mjd,
                      const ftk_search_params* params, long long* job, char* err, size_t err_len);
/* Waits for a job (timeout_s < 0: no limit): FTK_OK with its summary, FTK_ERROR with
 * the job's error, FTK_TIMEOUT. times (may be NULL, else 5 doubles): submitted,
 * admitted, ended (s on the server's clock), reserved bytes, device bytes its searcher
 * held at the end. */
int ftk_server_wait(ftk_server* server, long long job, double timeout_s, ftk_search_summary* summary, double* times,
                    char* err, size_t err_len);
/* Peaks of an ended job, as ftk_searcher_peaks; -1 for an unknown or failed job. */
int ftk_server_peaks(ftk_server* server, long long job, int which, int cap, int32_t* x, int32_t* y, float* snr,
                     int32_t* vidx);
/* Candidates of an ended job, as ftk_searcher_candidates; -1 for an unknown or failed job. */
int ftk_server_candidates(ftk_server* server, long long job, int cap, ftk_candidate* out);
/* Forgets an ended job. */
void ftk_server_release(ftk_server* server, long long job);
/* stats[13]: reserved bytes, peak reserved bytes, budget, idle-searcher bytes, running,
 * queued, completed, failed, evicted searchers, most jobs running at once, persistent
 * stacking launches, of those fused across jobs, passes stacked in fused launches. */
int ftk_server_stats(ftk_server* server, double* stats);
/* Cross-job fusion (default on, 3 ms): jobs of the same frame size whose persistent
 * stacking launches have the same work shape and meet within window_ms are stacked in
 * one launch (each job's result is unchanged). Jobs use the persistent launch unless
 * their params set queue = 0. */
int ftk_server_fusion(ftk_server* server, int on, double window_ms);

/* Plain device copy (n floats) on `stream`: the known-size kernel that calibrates
 * HBM traffic counters. */
int ftk_copy_f32(const float* d_src, float* d_dst, size_t count, void* stream);

/* Measured LDS read bandwidth of the current device (bytes/s) for reads of
 * width_bytes per lane: 4 (ds_read_b32), 8 (two dwords in one ds_read2st64_b32, the
 * form of the fp32 stacking loop) or 16 (ds_read_b128). The roofline of the stacking
 * kernel's inner loop, which reads one fp32 sample from LDS per stacked unit. */
int ftk_probe_lds(int width_bytes, double* bytes_per_s);

/* Host-only planning summary of a stacking pass (no device needed): shape[7] receives
 * the velocity blocks, the lattice blocks among them (4 x 4 motions sharing their
 * columns' dx and rows' dy in every frame, stacked by the lattice kernel), the motions
 * in lattice blocks, the largest per-frame vertical spread of a block, the segments
 * per tile, the blocks per segment and the motions not stacked because a lower-index
 * motion has the same shifts in every frame (identical sums: never the maximum). No
 * reference counterpart. */
int ftk_plan_shape(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count, int* shape,
                   char* err, size_t err_len);

/* Host emulation of the HIP algorithm (tiles, velocity blocks, -0.0 padding, edge
 * coverage counting, packed-key maximum) on the CPU, for tests without a GPU; the
 * analogue of emulate_cuda_pass (src/cuda_plan.cpp:74-117). frames: n x h x w
 * contiguous. */
int ftk_emulate_pass(int n, int w, int h, const float* frames, int nv, const int32_t* dx, const int32_t* dy,
                     int min_count, float* max_snr, int32_t* best_v, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* FTK_HIP_H */
