// Null-peak screen (is_artefact_peak, src/search.cpp:229-243) on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ftk {

// One pass's frames: element (x, y) of frame i at frames + i fstride + y pitch + x
// (fp32, or 16-bit times scale[i]), the pass's n frame times (hours) and, for a
// near-Earth search, its n parallax displacements at 1 AU (pixels; else null).
struct ScreenFrames {
    const void* frames;
    long long fstride;
    int pitch;
    int unused;
    const float* scale;
    const double* t;
    const double* par_x = nullptr;
    const double* par_y = nullptr;
};

// One launch screens the peaks of several passes: peak k belongs to passes[pslot[k]].
struct ScreenArgs {
    const ScreenFrames* passes = nullptr;  // device table
    const int* pslot = nullptr;
    int storage = 0;
    // background-subtracted frames (w x h each, sub_stride floats apart) and the static
    // sky: where either is non-finite the field is blank
    const float* sub = nullptr;
    long long sub_stride = 0;
    const float* sky = nullptr;
    const double* reg_dx = nullptr;  // n registration offsets each, or null (no detector test)
    const double* reg_dy = nullptr;
    const double* vel = nullptr;     // (vx, vy) per velocity index
    int nv = 0;                      // velocities; motion v is velocity v mod nv at inverse distance
    const double* qs = nullptr;      // qs[v / nv] (null: straight tracks, motion = velocity)
    const double* sqrtc = nullptr;   // sqrt(c) for c = 0 .. n
    const double* steps = nullptr;   // detector-test offsets: -span, -span + 0.5, ... <= span
    int nsteps = 0;
    const int* px = nullptr;         // peaks: grid position and velocity index
    const int* py = nullptr;
    const int* pv = nullptr;
    int count = 0;
    unsigned char* art = nullptr;    // out: 1 = artefact
    int n = 0, w = 0, h = 0;
    int edge_r = 0;                  // ceil(2 fwhm) + 1
    double max_frame_fraction = 0.25;
};

hipError_t launch_screen(const ScreenArgs& a, hipStream_t s);

}  // namespace ftk
