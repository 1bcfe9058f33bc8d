// Candidate refinement (refine, src/search.cpp:246-395) on the device: the per-peak
// searches and samplings of the real pass's frames, residuals, mask and subtracted
// frames; the host finishes each candidate from these figures.
#pragma once

#include <hip/hip_runtime.h>

namespace ftk {

// Figures of one refined candidate, kRefineOut doubles, in this order.
enum RefineOut {
    kRefX,        // refined track: position at t = 0 and motion (px, px/h)
    kRefY,
    kRefVx,
    kRefVy,
    kRefSnr,      // S/N of the refined track over all frames
    kRefHalf1,    // over the frames with t < 0, and with t >= 0
    kRefHalf2,
    kRefNum,      // PSF photometry sums over the stacked residual cutout: sum v k, sum k k
    kRefDen,
    kRefInside,   // frames where the track's sample is finite
    kRefMasked,   // of those, frames whose rounded track position is masked
    kRefEdge,     // 1: the field-edge test fires
    kRefDetector, // 1: the detector-fixed test fires
    kRefQ,        // refined inverse distance (1/AU; 0 on straight tracks)
    kRefOut
};

// One local-search level: motion step and half-range in steps, position step and
// half-range in steps.
struct RefineLevel {
    double vstep;
    double pstep;
    int nv;
    int np;
};

struct RefineArgs {
    const float* frames = nullptr;  // the real pass's filtered frames: n planes of w x h, fstride apart
    long long fstride = 0;
    const float* resid = nullptr;   // its residuals (frame - background - static sky), same layout
    const float* sub = nullptr;     // background-subtracted frames, same layout
    const float* sky = nullptr;     // static sky (w x h)
    const float* mask = nullptr;    // mask (w x h, 0 = masked)
    const double* t = nullptr;      // n frame times (h)
    const double* sqrtc = nullptr;  // sqrt(c), c = 0 .. n
    const double* steps = nullptr;  // detector-test offsets
    int nsteps = 0;
    const double* reg_dx = nullptr; // n registration offsets each, or null (no detector test)
    const double* reg_dy = nullptr;
    const double* kern = nullptr;   // photometry weights, (2 R + 1)^2, row-major
    int R = 0;                      // cutout half size
    const double* cand = nullptr;   // per candidate: x, y (grid position), vx, vy, q (grid motion)
    int count = 0;
    RefineLevel lev[2] = {};
    double dv = 0.0;                // velocity grid step (px/h)
    // near-Earth search: the frames' parallax displacements at 1 AU (n each) and the
    // inverse-distance step dq (> 0); null and 0 for straight tracks
    const double* par_x = nullptr;
    const double* par_y = nullptr;
    double dq = 0.0;
    int n = 0, w = 0, h = 0;
    int edge_r = 0;                 // ceil(2 fwhm) + 1
    double* out = nullptr;          // count x kRefOut
    float* contrib = nullptr;       // count x n: per-frame track samples (NaN where not finite)
};

hipError_t launch_refine(const RefineArgs& a, hipStream_t s);

}  // namespace ftk
