// C-ABI of the device preprocessing stage (include/ftk_hip.h, ftk_stage_*).
#include <hip/hip_runtime.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/ftk_hip.h"
#include "prep.hpp"

struct ftk_stage {
    ftk::DeviceStage* st = nullptr;
    int device = 0;
};

namespace {
void put_err(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    const size_t k = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = '\0';
}
}  // namespace

extern "C" int ftk_stage_create(int n, int w, int h, int device, ftk_stage** out, char* err, size_t err_len) {
    if (!out) return FTK_ERROR;
    *out = nullptr;
    try {
        if (device < 0) device = 0;
        if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("HIP device " + std::to_string(device) + " not present");
        auto* s = new ftk_stage();
        s->device = device;
        try {
            s->st = new ftk::DeviceStage(n, w, h);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" void ftk_stage_destroy(ftk_stage* s) {
    if (!s) return;
    delete s->st;
    delete s;
}

extern "C" int ftk_stage_background(ftk_stage* s, const float* d_raw, size_t raw_stride, float* d_sub,
                                    size_t sub_stride, double* background, void* stream, char* err, size_t err_len) {
    try {
        if (!s || !d_raw || !d_sub) throw std::runtime_error("ftk_stage_background: null argument");
        hipSetDevice(s->device);
        s->st->background(d_raw, static_cast<long long>(raw_stride), d_sub, static_cast<long long>(sub_stride),
                          background, static_cast<hipStream_t>(stream));
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_stage_build(ftk_stage* s, const float* d_sub, size_t sub_stride, const double* t_h,
                               const ftk_stage_params* p, float* d_filtered, size_t filt_stride, float* d_static_sky,
                               double* sigma, float* d_mask, int32_t* cr_pixels, void* stream, char* err,
                               size_t err_len) {
    try {
        if (!s || !d_sub || !t_h || !p || !d_filtered) throw std::runtime_error("ftk_stage_build: null argument");
        if (!(p->fwhm_px > 0.0)) throw std::runtime_error("ftk_stage_build: fwhm_px must be positive");
        hipSetDevice(s->device);
        ftk::StageParams q;
        q.fwhm_px = p->fwhm_px;
        q.clip_sigma = p->clip_sigma;
        q.cr_sigma = p->cr_sigma;
        q.static_exclude = p->static_exclude;
        q.mask_scatter = p->mask_scatter;
        q.weight_scatter = p->weight_scatter;
        s->st->build(d_sub, static_cast<long long>(sub_stride), t_h, q, d_filtered, static_cast<long long>(filt_stride),
                     d_static_sky, sigma, d_mask, cr_pixels, static_cast<hipStream_t>(stream));
        return FTK_OK;
    } catch (const std::exception& e) {
        put_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}
