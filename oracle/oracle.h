/* CPU oracle for the FaintTrail search hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP backend is
 * compared against (and the CPU baseline that bench.py times beside it). Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product library (fainttrail_amd/, libftk_hip.so) never links or calls it.
 *
 * It restates, in plain C++17 + OpenMP, the reference's algorithm for the path:
 *   - the stacking contract             include/fainttrail/backend.hpp:10-40
 *   - run_pass_naive / snr_at            src/backend.cpp:128-172
 *   - the cache-blocked CPU backend      src/backend_cpu.cpp:46-204
 *   - linear_shift_table (lround)        src/backend.cpp:114-123
 *   - check_backend                      src/backend.cpp:281-314
 *   - velocity_grid                      src/search.cpp:39-58
 *   - null_time_orders + xoshiro256**    src/search.cpp:559-571, src/random.hpp:14-85
 *   - build_stage (time-order stage)     src/preprocess.cpp:116-486
 *   - find_peaks                         src/search.cpp:105-135
 *   - set_threshold / tail fits          src/search.cpp:471-612
 * Every function cites the lines it follows. Pinning (how each function is checked
 * against the reference's own fixtures, the compiled reference in oracle/_ref, and
 * closed forms) is listed per function below; none is "parity unpinned".
 *
 * Frames are passed as one contiguous array, frame after frame: n x h x w fp32,
 * row-major, which is the layout the reference's Image vectors have one by one.
 */
#ifndef FTK_ORACLE_H
#define FTK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- stacking */

/* Spec-literal contract (backend.cpp:128-172): per pixel, per motion, per frame.
 * Pinned: known answer of test_main.cpp:439-444 (constant frames -> sqrt(7), index 0),
 * equality with oracle/_ref's run_pass_naive on the golden inputs. */
int ora_stack_naive(int n, int w, int h, const float* frames, int nv, const int32_t* dx, const int32_t* dy,
                    int min_count, float* max_snr, int32_t* best_v);

/* Cache-blocked OpenMP restatement of CpuBackend::run (backend_cpu.cpp:77-204):
 * bands of rows x blocks of motions, coverage from the frames' column intervals,
 * strict '>' in ascending motion order, non-finite maxima set to 0 at the end.
 * band_rows / vel_block <= 0 choose automatically (backend_cpu.cpp:46-65).
 * Pinned: bit-equal to ora_stack_naive for every blocking and thread count, and to
 * oracle/_ref's CpuBackend on the golden inputs. */
int ora_stack_blocked(int n, int w, int h, const float* frames, int nv, const int32_t* dx, const int32_t* dy,
                      int min_count, int threads, int band_rows, int vel_block, float* max_snr, int32_t* best_v);

/* Same, but only for the output rows [y0, y1) (a bounded CPU-baseline sample of a
 * large frame). frames holds the frame rows [fr0, fr1) of each frame (n x (fr1-fr0)
 * x w); coverage still uses the full h. Returns 2 when a row the outputs can reach
 * is missing. max_snr / best_v hold (y1 - y0) x w values. */
int ora_stack_blocked_rows(int n, int w, int h, const float* frames, int fr0, int fr1, int nv, const int32_t* dx,
                           const int32_t* dy, int min_count, int threads, int y0, int y1, float* max_snr,
                           int32_t* best_v);

/* Straight-line shifts dx = lround(vx t), dy = lround(vy t) in double
 * (backend.hpp:87-97, backend.cpp:114-123); motion-major j * n + i.
 * Pinned: closed form; a test compares with Python's exact rational rounding. */
void ora_linear_shifts(int n, const double* t_h, int nv, const double* vxy, int32_t* dx, int32_t* dy);

/* check_backend (backend.cpp:281-314): |a-b| <= 1e-4 max(1,|a|), NaN fails; a best_v
 * difference is explained only when both motions are defined there and their S/N
 * agree within the tolerance. out[0]=ok, out[1]=max_abs_diff, out[2]=bad values,
 * out[3]=best_v mismatches, out[4]=unexplained. Pinned: test_main.cpp:447-465
 * (+0.01 on one pixel and a wrong best_v are caught). */
void ora_check_backend(int n, int w, int h, const float* frames, int nv, const int32_t* dx, const int32_t* dy,
                       int min_count, const float* ref_max, const int32_t* ref_best, const float* cand_max,
                       const int32_t* cand_best, double* out);

/* ---------------------------------------------------------------- test data */

/* The reference tests' generator (test_main.cpp:28-36): xorshift64 (13, 7, 17),
 * uniform = ((s >> 11) + 0.5) / 2^53, normal by Box-Muller with the first draw in
 * the logarithm and the second in the cosine. Fills count floats. */
void ora_test_normals(uint64_t seed, size_t count, float* out);

/* xoshiro256** seeded by splitmix64 (random.hpp:14-85). */
typedef struct ora_rng { uint64_t s[4]; int have_spare; double spare; } ora_rng;
void ora_rng_seed(ora_rng* r, uint64_t seed);
uint64_t ora_rng_next(ora_rng* r);
double ora_rng_uniform(ora_rng* r);
double ora_rng_normal(ora_rng* r);

/* ---------------------------------------------------------------- search steps */

/* velocity_grid (search.cpp:39-58): vy-major, vx inner, -K..K, annulus test with
 * hypot. Writes up to cap pairs into vxy; returns the count (or -1 when the grid
 * exceeds 5e6 or the range is invalid). Pinned: closed form + oracle/_ref. */
int ora_velocity_grid(double span_h, double vmin_pxh, double vmax_pxh, double grid_tol_px, double* vxy, int cap,
                      double* dv_out);

/* null_time_orders (search.cpp:559-571): Fisher-Yates per run, one RNG stream.
 * out holds runs x n times. Pinned: golden permutation dump from oracle/_ref. */
void ora_null_time_orders(const double* t_h, int n, int runs, uint64_t seed, double* out);

/* find_peaks (search.cpp:105-135) without the final sort: row-major local maxima
 * (plateau rule strict to later, non-strict to earlier pixels) at or above floor.
 * Writes x, y, snr, vidx arrays (cap entries), returns the count. */
int ora_find_peaks_rowmajor(const float* max_snr, const int32_t* best_v, int w, int h, float floor, int radius,
                            int edge, int cap, int32_t* px, int32_t* py, float* psnr, int32_t* pv);

/* std::sort by descending snr of the row-major peak list, then truncate
 * (search.cpp:132-133). Same comparator, same libstdc++ algorithm. In place;
 * returns the kept count. */
int ora_sort_peaks(int count, int max_peaks, int32_t* px, int32_t* py, float* psnr, int32_t* pv);

/* Gaussian-tail null fit and threshold (search.cpp:480-527, 573-612 'gauss' branch
 * with the linear and quantile fallbacks and the floor). out[0]=threshold,
 * out[1]=a, out[2]=k, out[3]=inv_s2, out[4]=linear b, out[5]=method (0 gauss,
 * 1 linear, 2 quantile). */
void ora_null_threshold(const float* pooled, int count, int runs, double target_false_per_run,
                        double min_snr_floor, double* out);

/* ---------------------------------------------------------------- preprocessing */

typedef struct ora_stage_params {
    double fwhm_px;          /* PSF FWHM (estimated once on the real data)        */
    double clip_sigma;       /* 8                                                  */
    double cr_sigma;         /* 5                                                  */
    int static_exclude;      /* -1 = max(1, lround(n/8))                           */
    double mask_scatter;     /* 2.5                                                */
    double weight_scatter;   /* 1.5                                                */
    int threads;
} ora_stage_params;

/* The time-order-dependent stage build_stage (preprocess.cpp:286-486) for the
 * untiled case: static sky + leave-out residuals, per-frame sigma, scatter map,
 * mask, z, cosmic rays, clip, matched filter, normalisation.
 * sub: n x h x w background-subtracted frames; t_h: frame times (only their order
 * matters). Outputs: filtered (n x h x w), static_sky (h x w), sigma (n),
 * mask (h x w), cr_pixels (n). Any output pointer except filtered may be NULL. */
int ora_build_stage(int n, int w, int h, const float* sub, const double* t_h, const ora_stage_params* p,
                    float* filtered, float* static_sky, double* sigma, float* mask, int32_t* cr_pixels);

/* prepare_core's per-frame background and blank growth (preprocess.cpp:528-585,
 * 233-254): bg_i = median of the regular sample grid, sub = float(v - bg), NaN
 * areas grown by 3 px. frames n x h x w -> sub, background (n). */
int ora_background(int n, int w, int h, const float* frames, float* sub, double* background, int threads);

#ifdef __cplusplus
}
#endif

#endif
