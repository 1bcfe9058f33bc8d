// Device preprocessing (build_stage and the per-frame background) for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace ftk {

// One segment of a per-frame order-statistic selection: the regular sample grid of an
// image (sample_pixels, src/common.hpp:85-100) and the selection state.
struct SelectSeg {
    const float* img = nullptr;  // w x h image
    int w = 0;
    int border = 0, step = 1, nx = 0, ny = 0;
    int count = 0;               // valid samples
    int k[2] = {0, 0};           // remaining ranks of the two targets
    uint32_t prefix[2] = {0u, 0u};
    double center = 0.0;         // median (deviation mode centre)
    double result = 0.0;         // median, or robust sigma
};

// Parameters of build_stage (SearchParams, include/fainttrail/search.hpp:18-90).
struct StageParams {
    double fwhm_px = 3.0;
    double clip_sigma = 8.0;
    double cr_sigma = 5.0;
    int static_exclude = -1;
    double mask_scatter = 2.5;
    double weight_scatter = 1.5;
};

// Frames of one device stage: the static-sky kernel's per-pixel LDS columns (6 n bytes
// per lane beyond 64 frames) fit a workgroup's 160 KB of LDS up to this count.
constexpr int kMaxStageFrames = 420;

// Scratch and launch sequence for one frame stack of n x h x w.
class DeviceStage {
public:
    DeviceStage(int n, int w, int h);
    ~DeviceStage();
    DeviceStage(const DeviceStage&) = delete;
    DeviceStage& operator=(const DeviceStage&) = delete;

    // Per-frame background (median of the sample grid) and background-subtracted frames
    // with grown blank areas (prepare_core, src/preprocess.cpp:568-581).
    void background(const float* raw, long long raw_stride, float* sub, long long sub_stride, double* bg_host,
                    hipStream_t s);

    // build_stage for the frame times t_h (only their order matters): writes the
    // normalised matched-filtered frames; optional outputs may be null. resid_out
    // receives the residuals (frame - background - static sky, n planes of w x h).
    void build(const float* sub, long long sub_stride, const double* t_h, const StageParams& prm, float* filtered,
               long long filt_stride, float* static_sky, double* sigma_host, float* mask_out, int* cr_host,
               hipStream_t s, float* resid_out = nullptr);
    // The same with the time ranks already on the device (d_rank: n ints, time_ranks of
    // the frame times). Enqueues work only (no host copy) once the filter taps for
    // prm.fwhm_px are in place, so several builds can be queued on several streams.
    // defer_scale: the frames are left unnormalised, for copy_scaled to normalise while it
    // copies them into their stacking layout (one pass over the frames instead of two).
    void build_ranked(const float* sub, long long sub_stride, const int* d_rank, const StageParams& prm,
                      float* filtered, long long filt_stride, float* static_sky, double* sigma_host, float* mask_out,
                      int* cr_host, hipStream_t s, float* resid_out = nullptr, bool defer_scale = false);
    // The last build_ranked(defer_scale) frames (src, src_stride floats apart, w-float rows)
    // normalised into dst (frame i at dst + i dst_stride, rows dst_pitch apart), on s.
    void copy_scaled(const float* src, long long src_stride, float* dst, long long dst_stride, int dst_pitch,
                     hipStream_t s);
    // Matched-filter taps for prm.fwhm_px (uploaded when they change; synchronises s then).
    void set_taps(const StageParams& prm, hipStream_t s);
    // Device bytes of the stage's scratch (the sizes asked of hipMalloc).
    double device_bytes() const;

    int n, w, h;
    long long plane;

private:
    // Order-statistic selection over the sample grids (border) of the n images base + i
    // stride: median (deviation false) or robust sigma around the median of the
    // previous call (deviation true).
    void select(const float* base, long long stride, int border, const float* mask, bool deviation, int unit_floor,
                hipStream_t s);
    float* resid_ = nullptr;     // n planes: residuals, then z, then the row-filtered frames
    unsigned char* flags_ = nullptr;  // n planes of cosmic-ray hit flags / blank flags
    float* sky_ = nullptr;
    float* scatter_ = nullptr;
    float* mask_ = nullptr;
    SelectSeg* seg_ = nullptr;
    unsigned int* hist_ = nullptr;  // n x 2 x 256 digit counts, zero between passes
    float* inv_ = nullptr;
    double* sigma_ = nullptr;
    int* cr_ = nullptr;
    int* rank_ = nullptr;
    int* any_ = nullptr;
    float* taps_ = nullptr;
    double taps_sigma_ = -1.0;  // PSF sigma of the taps on the device
};

// Time ranks of n frame times: rank[i] = position of frame i in a stable sort by time.
void time_ranks(const double* t_h, int n, int* rank);

}  // namespace ftk
