// The "hip" SearchBackend for FaintTrail: the reference-side binding of the gfx950
// backend's C-ABI (include/ftk_hip.h).
//
// This file is written against the reference's own interface,
// include/fainttrail/backend.hpp:56-166 (SearchBackend, StackPass, StackResult,
// ShiftTable, register_backend, Cancelled), and is meant to be compiled into
// FaintTrail's library next to src/backend_cpu.cpp, linked with libftk_hip.so.
// Without any other change it registers itself under "hip" at start-up (the
// registry accepts that name, src/backend.cpp:27-35), so
//   fainttrail search DIR --backend hip
//   fainttrail check-backend --backend hip
//   ft_search() with ft_params.backend = "hip"
// run unchanged. Options come from FAINTTRAIL_HIP_OPTIONS ("device=0,max_mb=2,...",
// the same keys as "cuda:...", src/cuda_plan.cpp:14-33), since registered names are
// matched whole.
//
// Built with FAINTTRAIL_HAVE_HIP (the reference's backend.cpp then knows the backend as
// it knows CUDA: "hip" and "hip:key=value,..." names, "auto" preferring it, listed
// first by backend_names; see oracle/hip_integration.sh for the edits), it does not
// register itself and exports hip_backend_available() and make_hip_backend() instead.
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fainttrail/backend.hpp"
#include "ftk_hip.h"

namespace fainttrail::detail {

namespace {

class HipBackend final : public SearchBackend {
public:
    explicit HipBackend(const std::string& spec) {
        char err[512] = {0};
        if (ftk_parse_options(spec.c_str(), &opt_, err, sizeof err) != FTK_OK) throw std::runtime_error(err);
        char why[512] = {0};
        if (ftk_device_count(why, sizeof why) <= 0) throw std::runtime_error(std::string("HIP backend unavailable: ") + why);
        char d[512] = {0};
        ftk_device_description(opt_.device, d, sizeof d);
        desc_ = d;
    }
    std::string name() const override { return "hip"; }
    std::string description() const override { return desc_; }

    StackResult run(const StackPass& pass, const ProgressFn& progress) override {
        validate_pass(pass);
        const auto& frames = *pass.frames;
        const int n = static_cast<int>(frames.size()), w = frames[0].width, h = frames[0].height;
        // the contract's straight-line rule and an explicit table are the same thing
        // for the backend (linear_shift_table, src/backend.cpp:114-123)
        ShiftTable lin;
        const ShiftTable* tab = pass.shifts;
        if (!tab) {
            lin = linear_shift_table(*pass.t_h, *pass.velocities);
            tab = &lin;
        }
        std::vector<const float*> ptrs(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) ptrs[static_cast<size_t>(i)] = frames[static_cast<size_t>(i)].data.data();
        ftk_pass p{};
        p.n = n;
        p.w = w;
        p.h = h;
        p.frames = ptrs.data();
        p.nv = tab->motions;
        p.dx = tab->dx.data();
        p.dy = tab->dy.data();
        p.min_count = pass.min_count;
        StackResult out;
        out.max_snr = Image(w, h, 0.0f);
        out.best_v.assign(static_cast<size_t>(w) * h, -1);
        ftk_timings t{};
        struct Ctx {
            const ProgressFn* fn;
        } ctx{&progress};
        auto cb = [](double f, const char* stage, void* user) -> int {
            const auto* c = static_cast<const Ctx*>(user);
            return (*c->fn)(f, stage) ? 1 : 0;
        };
        char err[1024] = {0};
        const int st = ftk_stack_pass(&p, &opt_, out.max_snr.data.data(), out.best_v.data(), &t,
                                      progress ? +cb : nullptr, &ctx, err, sizeof err);
        if (st == FTK_CANCELLED) throw Cancelled();
        if (st != FTK_OK) throw std::runtime_error(err);
        out.timings.push_back({"upload", t.upload_s});
        out.timings.push_back({"kernel", t.kernel_s});
        out.timings.push_back({"download", t.download_s});
        out.timings.push_back({"setup", t.setup_s});
        return out;
    }

private:
    ftk_options opt_{};
    std::string desc_;
};

#ifndef FAINTTRAIL_HAVE_HIP
struct Registration {
    Registration() {
        register_backend("hip", [] {
            const char* spec = std::getenv("FAINTTRAIL_HIP_OPTIONS");
            return std::unique_ptr<SearchBackend>(new HipBackend(spec ? spec : ""));
        });
    }
};
const Registration registration;
#endif

}  // namespace

// The built-in form's entry points (the counterparts of cuda_device_count and
// make_cuda_backend, src/backend_cuda.cu:198-209, 80-97).
bool hip_backend_available(std::string* why) {
    char buf[512] = {0};
    const bool ok = ftk_device_count(buf, sizeof buf) > 0;
    if (!ok && why) *why = buf;
    return ok;
}

std::unique_ptr<SearchBackend> make_hip_backend(const std::string& options) {
    // the options after "hip:", else FAINTTRAIL_HIP_OPTIONS
    const char* env = std::getenv("FAINTTRAIL_HIP_OPTIONS");
    return std::unique_ptr<SearchBackend>(new HipBackend(!options.empty() ? options : (env ? env : "")));
}

}  // namespace fainttrail::detail
