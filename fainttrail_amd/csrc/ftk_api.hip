// C-ABI of the gfx950 stacking backend (include/ftk_hip.h): options, devices, the
// synchronous drop-in pass (host frames in, host maps out), the device-resident plan,
// and the host emulation of the kernel's algorithm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ftk_hip.h"
#include "stack_launch.hpp"
#include "store.hpp"
#include "stack_plan.hpp"

namespace ftk {

namespace {

void set_err(char* err, size_t len, const std::string& msg) {
    if (!err || len == 0) return;
    const size_t k = std::min(len - 1, msg.size());
    std::memcpy(err, msg.data(), k);
    err[k] = '\0';
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

struct Cancel {};

// Launches of a host-buffer pass, alternating between two streams so that one launch's
// last workgroups overlap the next launch's first (launches write disjoint output rows,
// or merge into the keys by atomic maximum), with an event per launch and up to three
// launches queued: progress is reported as each launch completes, in order. Destruction (normal or on cancel or
// error) waits for everything launched.
struct LaunchQueue {
    hipStream_t st[2] = {nullptr, nullptr};
    std::deque<std::pair<hipEvent_t, double>> pending;  // launch done, fraction of the pass after it
    int count = 0;
    LaunchQueue() {
        for (auto& s : st) check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "create stream");
    }
    ~LaunchQueue() {
        for (auto& e : pending) {
            (void)hipEventSynchronize(e.first);
            (void)hipEventDestroy(e.first);
        }
        for (auto& s : st)
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
    }
    hipStream_t next() { return st[count++ & 1]; }
    void launched(hipStream_t s, double done) {
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "create event");
        pending.emplace_back(e, done);
        check(hipEventRecord(e, s), "record event");
    }
    // waits for launches until at most `keep` are in flight; false: the progress callback cancelled
    bool settle(size_t keep, ftk_progress_fn progress, void* user) {
        while (pending.size() > keep) {
            auto [e, done] = pending.front();
            pending.pop_front();
            const hipError_t r = hipEventSynchronize(e);
            (void)hipEventDestroy(e);
            check(r, "stack kernel");
            if (progress && !progress(done, "stack", user)) return false;
        }
        return true;
    }
};

template <typename T>
struct DeviceBuffer {
    T* p = nullptr;
    size_t n = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
    void reserve(size_t count, const char* what) {
        if (count <= n && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        check(hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(1, count) * sizeof(T)), what);
        n = count;
    }
    void upload(const std::vector<T>& v, const char* what) {
        reserve(v.size(), what);
        if (v.empty()) return;
        check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), what);
        // a copy from pageable memory may return before its DMA has landed on the null
        // stream, which the launches' non-blocking streams do not wait for
        check(hipStreamSynchronize(nullptr), what);
    }
};

int resolve_device(int requested) {
    if (requested >= 0) return requested;
    if (const char* e = std::getenv("FAINTTRAIL_HIP_DEVICE")) return std::atoi(e);
    return 0;
}

ftk_options default_options() {
    ftk_options o;
    o.device = -1;
    o.max_mb = 0.0;
    // about 0.25 s of kernel time per launch at the measured stacking rate: progress and
    // cancel stay responsive, and launch boundaries cost a few % (1e11, 10 ms launches,
    // cost ~20 % on the 6144 x 4096 x 60 x 19952 pass when each launch was waited for)
    o.launch_adds = 4e12;
    o.pad = 1;
    o.tile_rows = 0;
    o.queue = -1;
    return o;
}

// The plan's tiles at a work shape.
void use_shape(StackPlan& p, const StackShape& sh, int out_rows) {
    if (p.tile_rows != sh.tile_rows || p.tiles_y == 0) {
        p.tile_rows = sh.tile_rows;
        plan_tiles(p, out_rows);
    }
}

// Plan tables on the device.
struct DevicePlan {
    DeviceBuffer<int> shift, box, ubox, lds_off, motion, cov_x, cov_y, chain, lat;
    DeviceBuffer<long long> cov_off;
    DeviceBuffer<float> inv_sqrt;
    void upload(const StackPlan& p) {
        shift.upload(p.shift, "upload shift table");
        box.upload(p.box, "upload boxes");
        ubox.upload(p.ubox, "upload boxes");
        lds_off.upload(p.lds_off, "upload offsets");
        cov_off.upload(p.cov_off, "upload coverage tables");
        cov_x.upload(p.cov_x, "upload coverage tables");
        cov_y.upload(p.cov_y, "upload coverage tables");
        chain.upload(p.chain, "upload coverage tables");
        lat.upload(p.lat, "upload lattice tables");
        motion.upload(p.motion, "upload motions");
        inv_sqrt.upload(p.inv_sqrt, "upload 1/sqrt table");
    }
};

}  // namespace
}  // namespace ftk

using namespace ftk;

struct ftk_plan {
    StackPlan plan;
    int device = 0;
    DevicePlan dev;
    DeviceBuffer<unsigned long long> keys;
    // the frames in a zero-bordered buffer (every staged box inside it), when that
    // costs at most twice the frames' own size
    bool use_pad = false;
    PadGeometry pad;
    DeviceBuffer<float> padded;
    // 16-bit storage (ftk_plan_store): frames in a zero-bordered 16-bit buffer with a
    // power-of-two scale per frame
    int storage = 0;
    PadGeometry pad16;
    DeviceBuffer<uint16_t> stored;
    DeviceBuffer<float> scales;
    DeviceBuffer<unsigned> maxbits;
    // work shapes of fp32 runs and of 16-bit runs (from the options, or ftk_plan_tune)
    // and the asked values (tile_rows 0, queue -1: automatic)
    StackShape shape32, shape16;
    int ask_tile_rows = 0, ask_queue = -1;
    // the persistent launch's counters and descriptors
    DeviceBuffer<int> queue_ws;
    // otherwise the edge tiles' launch runs on a stream of its own beside the interior launch
    hipStream_t ring_stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    ~ftk_plan() {
        if (ring_stream) (void)hipStreamSynchronize(ring_stream);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (ring_stream) (void)hipStreamDestroy(ring_stream);
    }
};

namespace ftk {
namespace {
// Whether a padded layout is worth its memory: at most twice the frames' own size.
bool pad_worth(const PadGeometry& g, int w, int rows) {
    return static_cast<double>(g.pitch) * g.rows <= 2.0 * static_cast<double>(w) * rows;
}
}  // namespace
}  // namespace ftk

extern "C" int ftk_parse_options(const char* spec, ftk_options* out, char* err, size_t err_len) {
    if (!out) return FTK_ERROR;
    ftk_options o = default_options();
    try {
        std::stringstream ss(spec ? spec : "");
        std::string item;
        while (std::getline(ss, item, ',')) {
            if (item.empty()) continue;
            const auto eq = item.find('=');
            const std::string key = item.substr(0, eq), val = eq == std::string::npos ? "" : item.substr(eq + 1);
            char* end = nullptr;
            const double d = std::strtod(val.c_str(), &end);
            if (val.empty() || end == val.c_str() || *end != '\0')
                throw std::runtime_error("hip backend option '" + item + "': expected key=number");
            if (key == "device") o.device = static_cast<int>(d);
            else if (key == "max_mb") o.max_mb = d;
            else if (key == "launch_adds") o.launch_adds = d;
            else if (key == "pad" && (d == 0.0 || d == 1.0)) o.pad = static_cast<int>(d);
            else if (key == "pad") throw std::runtime_error("hip backend option 'pad' must be 0 or 1");
            else if (key == "tile_rows" && (d == 0.0 || (tile_rows_ok(static_cast<int>(d)) && d == static_cast<int>(d))))
                o.tile_rows = static_cast<int>(d);
            else if (key == "tile_rows") throw std::runtime_error("hip backend option 'tile_rows' must be 0, 8, 16 or 32");
            else if (key == "queue" && (d == -1.0 || d == 0.0 || d == 1.0)) o.queue = static_cast<int>(d);
            else if (key == "queue") throw std::runtime_error("hip backend option 'queue' must be -1, 0 or 1");
            else throw std::runtime_error("unknown hip backend option '" + key + "' (device, max_mb, launch_adds, pad, tile_rows, queue)");
        }
        if (o.max_mb < 0 || o.launch_adds <= 0) throw std::runtime_error("hip backend options must be positive");
        if (o.queue == 0 && o.tile_rows != 0 && o.tile_rows != kTileY)
            throw std::runtime_error("hip backend option 'tile_rows' other than 16 needs queue=1");
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
    *out = o;
    return FTK_OK;
}

extern "C" int ftk_device_count(char* why, size_t why_len) {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) {
        set_err(why, why_len, hipGetErrorString(e));
        (void)hipGetLastError();
        return 0;
    }
    if (count == 0) set_err(why, why_len, "no HIP device found");
    return count;
}

extern "C" int ftk_device_description(int device, char* buf, size_t len) {
    device = resolve_device(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        set_err(buf, len, "HIP device " + std::to_string(device) + " not present");
        return FTK_ERROR;
    }
    char tmp[512];
    std::snprintf(tmp, sizeof tmp, "HIP device %d: %s (%s), %.1f GB, %d CUs", device, prop.name, prop.gcnArchName,
                  prop.totalGlobalMem / 1073741824.0, prop.multiProcessorCount);
    set_err(buf, len, tmp);
    return FTK_OK;
}

extern "C" int ftk_stack_pass(const ftk_pass* pass, const ftk_options* opt_in, float* max_snr, int32_t* best_v,
                              ftk_timings* timings, ftk_progress_fn progress, void* user, char* err, size_t err_len) {
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    try {
        if (!pass || !pass->frames) throw std::runtime_error("stack pass: missing frames");
        if (!max_snr || !best_v) throw std::runtime_error("stack pass: missing outputs");
        for (int i = 0; i < pass->n; ++i)
            if (!pass->frames[i]) throw std::runtime_error("stack pass: missing frames");
        const ftk_options opt = opt_in ? *opt_in : default_options();
        StackPlan p = make_stack_plan(pass->n, pass->w, pass->h, pass->nv, pass->dx, pass->dy, pass->min_count);
        const StackShape shape = resolve_shape(p, kStoreF32, opt.tile_rows, opt.queue, p.h);
        use_shape(p, shape, p.h);
        const int n = p.n, w = p.w, h = p.h, ty = p.tile_rows;
        const int device = resolve_device(opt.device);
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count)
            throw std::runtime_error("HIP device " + std::to_string(device) + " not present (" + std::to_string(count) + " found)");
        check(hipSetDevice(device), "hipSetDevice");

        // memory: the frame rows of a band take the budget; tables are small and fixed
        size_t free_b = 0, total_b = 0;
        check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const double fixed = 4.0 * (p.shift.size() + p.box.size() + p.motion.size()) + 64.0 * 1048576.0;
        double budget = 0.8 * static_cast<double>(free_b) - fixed;
        if (opt.max_mb > 0) budget = std::min(budget, opt.max_mb * 1048576.0);
        // per output row a band also holds keys (8 w B) and the two output maps (8 w B)
        const double share = static_cast<double>(n) / (n + 4.0);
        // rows are budgeted at the padded width (a zero border for the shifts, see below)
        const int w_pad = std::max(0, -p.all_box[0]) + std::max(w, kTileX * p.tiles_x + std::max(0, p.all_box[2]));
        const std::vector<Band> bands = plan_row_bands(w_pad, h, n, p.min_dy, p.max_dy, std::max(1.0, budget * share));

        DevicePlan dev;
        dev.upload(p);
        DeviceBuffer<float> d_frames, d_max, d_stage;
        DeviceBuffer<int> d_best;
        // persistent launches: a counter workspace per launch slot; launches cycle through
        // 8 slots and at most 3 are in flight (LaunchQueue::settle(2) below)
        constexpr int kSlots = 8;
        constexpr size_t kSlotInts = kQueueWorkspaceBytes / sizeof(int);
        DeviceBuffer<int> d_queue;
        if (shape.queue) d_queue.reserve(kSlots * kSlotInts, "allocate queue counters");
        int launch_no = 0;
        DeviceBuffer<unsigned long long> d_keys;
        double t_up = 0, t_kernel = 0, t_down = 0;
        double done = 0;  // fraction of the output rows stacked
        for (const Band& b : bands) {
            const int rows = b.r1 - b.r0, out_rows = b.y1 - b.y0;
            const size_t np = static_cast<size_t>(out_rows) * w;
            const double t_up_band0 = t_up;
            plan_tiles(p, out_rows);
            // the band's frame rows inside a zero border that holds every staged box, when
            // that costs at most twice the rows' own size; else plain rows
            const PadGeometry g = pad_geometry(p, b.r0, b.r1, b.y0, p.tiles_x, p.tiles_y);
            const bool use_pad = opt.pad != 0 && pad_worth(g, w, rows);
            const long long stride = use_pad ? static_cast<long long>(g.rows) * g.pitch : static_cast<long long>(rows) * w;
            const long long origin = use_pad ? static_cast<long long>(g.top) * g.pitch + g.left : 0;
            d_frames.reserve(static_cast<size_t>(n) * stride, "allocate frames");
            d_keys.reserve(np, "allocate keys");
            d_max.reserve(np, "allocate output");
            d_best.reserve(np, "allocate output");
            // Frame rows go up in row chunks, each just before the first launch that reads
            // them, so the stacking of the rows above runs while the rest uploads: every
            // frame's rows of a chunk as one contiguous copy (pitched copies from pageable
            // memory go row by row), into the zero border through a device-side row copy.
            // The copies are synchronous on the null stream; the launches run on
            // non-blocking streams, so they keep running meanwhile.
            const size_t chunk_cap = static_cast<size_t>(rows) * w;
            if (use_pad) {
                check(hipMemset(d_frames.p, 0, sizeof(float) * n * stride), "clear frames");
                check(hipStreamSynchronize(nullptr), "clear frames");  // before non-blocking streams read it
                d_stage.reserve(chunk_cap, "allocate upload buffer");
            }
            int up_to = b.r0;  // frame rows [b.r0, up_to) are on the device
            auto upload_until = [&](int r) {
                r = std::min(r, b.r1);
                if (r <= up_to) return;
                const auto u0 = clk::now();
                const size_t cnt = static_cast<size_t>(r - up_to) * w;
                for (int i = 0; i < n; ++i) {
                    const float* src = pass->frames[i] + static_cast<size_t>(up_to) * w;
                    if (use_pad) {
                        // the previous row copy has read the stage buffer (same stream)
                        check(hipMemcpy(d_stage.p, src, sizeof(float) * cnt, hipMemcpyHostToDevice), "upload frames");
                        check(launch_copy_rows(d_stage.p, static_cast<long long>(cnt),
                                               d_frames.p + i * stride + origin + static_cast<long long>(up_to - b.r0) * g.pitch,
                                               stride, g.pitch, w, r - up_to, 1, nullptr),
                              "pad frames");
                    } else {
                        check(hipMemcpy(d_frames.p + i * stride + static_cast<long long>(up_to - b.r0) * w, src,
                                        sizeof(float) * cnt, hipMemcpyHostToDevice),
                              "upload frames");
                    }
                }
                // the row copies (and the direct uploads, whose DMA may land after
                // hipMemcpy returns) are on the null stream: complete before the launches
                check(hipStreamSynchronize(nullptr), "upload frames");
                up_to = r;
                t_up += std::chrono::duration<double>(clk::now() - u0).count();
            };
            auto t1 = clk::now();

            // launches of about launch_adds units: whole tile rows of the band with every
            // velocity block (all segments of a tile in one launch share its frame boxes
            // in L2); when one tile row is already more, segments of its blocks
            const double units_per_row = static_cast<double>(w) * n * p.nv;
            int rows_per_launch = static_cast<int>(std::clamp(opt.launch_adds / std::max(1.0, units_per_row), 1.0,
                                                              static_cast<double>(out_rows)));
            rows_per_launch = std::max(ty, rows_per_launch / ty * ty);
            check(launch_fill_keys(d_keys.p, static_cast<long long>(np), nullptr), "clear keys");
            check(hipDeviceSynchronize(), "clear keys");
            LaunchQueue lq;
            for (int y = b.y0; y < b.y1; y += rows_per_launch) {
                const int sub_rows = std::min(rows_per_launch, b.y1 - y);
                // the launch's tiles read frame rows up to (its last tile row) + max_dy
                upload_until(y + (sub_rows + ty - 1) / ty * ty + p.max_dy);
                plan_tiles(p, sub_rows);
                const double units_per_block = static_cast<double>(sub_rows) * w * n * kVel;
                const int blocks_per_launch = static_cast<int>(
                    std::clamp(opt.launch_adds / std::max(1.0, units_per_block), 1.0, static_cast<double>(p.nb)));
                const int bps = std::min(p.blocks_per_seg, blocks_per_launch);
                const int nseg_total = (p.nb + bps - 1) / bps;
                const int segs_per_launch = std::max(1, blocks_per_launch / bps);
                for (int s0 = 0; s0 < nseg_total; s0 += segs_per_launch) {
                    StackLaunch L{};
                    L.frames = d_frames.p + origin;
                    L.frame_stride = stride;
                    L.pitch = use_pad ? g.pitch : w;
                    L.padded = use_pad;
                    L.row0 = b.r0;
                    L.row1 = b.r1;
                    L.shift = dev.shift.p;
                    L.box = dev.box.p;
                    L.ubox = dev.ubox.p;
                    L.lds_off = dev.lds_off.p;
                    L.cov_off = dev.cov_off.p;
                    L.cov_x = dev.cov_x.p;
                    L.cov_y = dev.cov_y.p;
                    L.chain = dev.chain.p;
                    L.lat = p.nb_lat > 0 ? dev.lat.p : nullptr;
                    L.motion = dev.motion.p;
                    L.inv_sqrt = dev.inv_sqrt.p;
                    L.y_begin = y;
                    L.y_end = y + sub_rows;
                    L.tiles_x = p.tiles_x;
                    L.tiles_y = p.tiles_y;
                    L.bps = bps;
                    L.seg0 = s0;
                    L.nseg = std::min(segs_per_launch, nseg_total - s0);
                    L.keys = d_keys.p + static_cast<size_t>(y - b.y0) * w;
                    if (shape.queue) L.counters = d_queue.p + (launch_no++ % kSlots) * kSlotInts;
                    hipStream_t st = lq.next();
                    check(launch_stack(p, L, st), "launch stack kernel");
                    const int blocks_done = std::min(p.nb, (s0 + L.nseg) * bps) - std::min(p.nb, s0 * bps);
                    done += static_cast<double>(sub_rows) / h * blocks_done / p.nb;
                    lq.launched(st, done);
                    if (!lq.settle(2, progress, user)) throw Cancel{};
                }
            }
            if (!lq.settle(0, progress, user)) throw Cancel{};
            check(launch_finalize(d_keys.p, static_cast<long long>(np), d_max.p, d_best.p, nullptr), "finalize");
            check(hipDeviceSynchronize(), "finalize");
            auto t2 = clk::now();
            // stacking time net of the uploads it overlapped
            t_kernel += std::chrono::duration<double>(t2 - t1).count() - (t_up - t_up_band0);
            check(hipMemcpy(max_snr + static_cast<size_t>(b.y0) * w, d_max.p, np * sizeof(float), hipMemcpyDeviceToHost),
                  "download");
            check(hipMemcpy(best_v + static_cast<size_t>(b.y0) * w, d_best.p, np * sizeof(int), hipMemcpyDeviceToHost),
                  "download");
            t_down += std::chrono::duration<double>(clk::now() - t2).count();
        }
        const double total = std::chrono::duration<double>(clk::now() - t_start).count();
        if (timings) {
            timings->upload_s = t_up;
            timings->kernel_s = t_kernel;
            timings->download_s = t_down;
            timings->setup_s = std::max(0.0, total - t_up - t_kernel - t_down);
        }
        if (progress) progress(1.0, "stack", user);
        return FTK_OK;
    } catch (const Cancel&) {
        set_err(err, err_len, "cancelled");
        return FTK_CANCELLED;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_plan_create(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count,
                               const ftk_options* opt_in, ftk_plan** out, char* err, size_t err_len) {
    if (!out) return FTK_ERROR;
    *out = nullptr;
    try {
        const ftk_options opt = opt_in ? *opt_in : default_options();
        auto* pl = new ftk_plan();
        try {
            pl->plan = make_stack_plan(n, w, h, nv, dx, dy, min_count);
            pl->ask_tile_rows = opt.tile_rows;
            pl->ask_queue = opt.queue;
            pl->shape32 = resolve_shape(pl->plan, kStoreF32, opt.tile_rows, opt.queue, h);
            use_shape(pl->plan, pl->shape32, h);
            pl->device = resolve_device(opt.device);
            check(hipSetDevice(pl->device), "hipSetDevice");
            pl->dev.upload(pl->plan);
            pl->queue_ws.reserve(kQueueWorkspaceBytes / sizeof(int), "allocate queue counters");
            pl->keys.reserve(static_cast<size_t>(w) * h, "allocate keys");
            const StackPlan& p = pl->plan;
            pl->pad = pad_geometry(p, 0, h, 0, p.tiles_x, p.tiles_y);
            pl->use_pad = opt.pad != 0 && pad_worth(pl->pad, w, h);
            if (pl->use_pad) {
                const size_t count = static_cast<size_t>(n) * pl->pad.rows * pl->pad.pitch;
                pl->padded.reserve(count, "allocate padded frames");
                check(hipMemset(pl->padded.p, 0, count * sizeof(float)), "clear padded frames");
                check(hipStreamSynchronize(nullptr), "clear padded frames");  // before non-blocking streams read it
            }
        } catch (...) {
            delete pl;
            throw;
        }
        *out = pl;
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

namespace ftk {
namespace {
// The launch of a whole-frame pass over the plan's tables; the caller sets the frames.
StackLaunch plan_launch(ftk_plan* pl) {
    const StackPlan& p = pl->plan;
    StackLaunch L{};
    L.row0 = 0;
    L.row1 = p.h;
    L.shift = pl->dev.shift.p;
    L.box = pl->dev.box.p;
    L.ubox = pl->dev.ubox.p;
    L.lds_off = pl->dev.lds_off.p;
    L.cov_off = pl->dev.cov_off.p;
    L.cov_x = pl->dev.cov_x.p;
    L.cov_y = pl->dev.cov_y.p;
    L.chain = pl->dev.chain.p;
    L.lat = p.nb_lat > 0 ? pl->dev.lat.p : nullptr;
    L.motion = pl->dev.motion.p;
    L.inv_sqrt = pl->dev.inv_sqrt.p;
    L.y_begin = 0;
    L.y_end = p.h;
    L.tiles_x = p.tiles_x;
    L.tiles_y = p.tiles_y;
    L.bps = p.blocks_per_seg;
    L.seg0 = 0;
    L.nseg = p.nseg;
    L.keys = pl->keys.p;
    return L;
}

// One pass of the plan at its shape for the launch's storage (the caller built L with
// plan_launch after use_shape for that shape).
void run_launch(ftk_plan* pl, const StackLaunch& L_in, float* d_max_snr, int32_t* d_best_v, hipStream_t s) {
    const long long np = static_cast<long long>(pl->plan.w) * pl->plan.h;
    const StackShape& sh = L_in.storage == kStoreF32 ? pl->shape32 : pl->shape16;
    if (sh.queue) {
        StackLaunch L = L_in;
        L.counters = pl->queue_ws.p;
        check(launch_fill_keys(pl->keys.p, np, s), "clear keys");
        check(launch_stack(pl->plan, L, s), "launch stack kernel");
        check(launch_finalize(pl->keys.p, np, d_max_snr, d_best_v, s), "finalize");
        return;
    }
    const StackLaunch& L = L_in;
    if (!pl->ring_stream) {
        check(hipStreamCreateWithFlags(&pl->ring_stream, hipStreamNonBlocking), "create stream");
        check(hipEventCreateWithFlags(&pl->ev_in, hipEventDisableTiming), "create event");
        check(hipEventCreateWithFlags(&pl->ev_out, hipEventDisableTiming), "create event");
    }
    check(launch_fill_keys(pl->keys.p, np, s), "clear keys");
    check(hipEventRecord(pl->ev_in, s), "record event");
    check(hipStreamWaitEvent(pl->ring_stream, pl->ev_in, 0), "stream wait");
    check(launch_stack(pl->plan, L, s, pl->ring_stream), "launch stack kernel");
    check(hipEventRecord(pl->ev_out, pl->ring_stream), "record event");
    check(hipStreamWaitEvent(s, pl->ev_out, 0), "stream wait");
    check(launch_finalize(pl->keys.p, np, d_max_snr, d_best_v, s), "finalize");
}
}  // namespace
}  // namespace ftk

extern "C" int ftk_plan_run(ftk_plan* pl, const float* d_frames, size_t frame_stride, float* d_max_snr,
                            int32_t* d_best_v, void* stream, char* err, size_t err_len) {
    try {
        if (!pl || !d_frames || !d_max_snr || !d_best_v) throw std::runtime_error("ftk_plan_run: null argument");
        const StackPlan& p = pl->plan;
        if (frame_stride < static_cast<size_t>(p.w) * p.h) throw std::runtime_error("ftk_plan_run: frame stride smaller than a frame");
        check(hipSetDevice(pl->device), "hipSetDevice");
        hipStream_t s = static_cast<hipStream_t>(stream);
        use_shape(pl->plan, pl->shape32, p.h);
        StackLaunch L = plan_launch(pl);
        L.frames = d_frames;
        L.frame_stride = static_cast<long long>(frame_stride);
        if (pl->use_pad) {
            // the frames into the interior of the zero-bordered buffer (its border stays 0)
            const PadGeometry& g = pl->pad;
            const long long stride = static_cast<long long>(g.rows) * g.pitch;
            float* origin = pl->padded.p + static_cast<long long>(g.top) * g.pitch + g.left;
            check(launch_copy_rows(d_frames, static_cast<long long>(frame_stride), origin, stride, g.pitch, p.w, p.h,
                                   p.n, s),
                  "pad frames");
            L.frames = origin;
            L.frame_stride = stride;
            L.pitch = g.pitch;
            L.padded = true;
        }
        run_launch(pl, L, d_max_snr, d_best_v, s);
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_plan_store(ftk_plan* pl, const float* d_frames, size_t frame_stride, int storage, void* stream,
                              char* err, size_t err_len) {
    try {
        if (!pl || !d_frames) throw std::runtime_error("ftk_plan_store: null argument");
        if (storage != FTK_STORE_F16 && storage != FTK_STORE_BF16)
            throw std::runtime_error("ftk_plan_store: storage must be FTK_STORE_F16 or FTK_STORE_BF16");
        const StackPlan& p = pl->plan;
        if (frame_stride < static_cast<size_t>(p.w) * p.h) throw std::runtime_error("ftk_plan_store: frame stride smaller than a frame");
        check(hipSetDevice(pl->device), "hipSetDevice");
        hipStream_t s = static_cast<hipStream_t>(stream);
        // the 16-bit runs' shape (a tile height the 16-bit staging cannot hold for this
        // plan becomes 16 rows; the padded layout is the same for every tile height)
        if (!pl->storage || pl->storage != storage)
            pl->shape16 = resolve_shape(p, storage, pl->ask_tile_rows, pl->ask_queue, p.h);
        if (!pl->stored.p) {
            // 16-bit frames are always padded (their fetch reads whole pixel pairs)
            pl->pad16 = pad_geometry16(pad_geometry(p, 0, p.h, 0, p.tiles_x, p.tiles_y));
            const size_t count = static_cast<size_t>(p.n) * pl->pad16.rows * pl->pad16.pitch;
            pl->stored.reserve(count, "allocate 16-bit frames");
            check(hipMemsetAsync(pl->stored.p, 0, count * sizeof(uint16_t), s), "clear 16-bit frames");
            pl->scales.reserve(static_cast<size_t>(p.n), "allocate scales");
            pl->maxbits.reserve(static_cast<size_t>(p.n), "allocate scratch");
        }
        const PadGeometry& g = pl->pad16;
        const long long stride = static_cast<long long>(g.rows) * g.pitch;
        uint16_t* origin = pl->stored.p + static_cast<long long>(g.top) * g.pitch + g.left;
        check(launch_store(d_frames, static_cast<long long>(frame_stride), p.w, p.n, p.w, p.h, storage, origin, stride,
                           g.pitch, pl->maxbits.p, pl->scales.p, s),
              "store frames");
        pl->storage = storage;
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_plan_run_stored(ftk_plan* pl, float* d_max_snr, int32_t* d_best_v, void* stream, char* err,
                                   size_t err_len) {
    try {
        if (!pl || !d_max_snr || !d_best_v) throw std::runtime_error("ftk_plan_run_stored: null argument");
        if (!pl->storage) throw std::runtime_error("ftk_plan_run_stored: no frames stored (ftk_plan_store)");
        check(hipSetDevice(pl->device), "hipSetDevice");
        const PadGeometry& g = pl->pad16;
        use_shape(pl->plan, pl->shape16, pl->plan.h);
        StackLaunch L = plan_launch(pl);
        L.frames = pl->stored.p + static_cast<long long>(g.top) * g.pitch + g.left;
        L.frame_stride = static_cast<long long>(g.rows) * g.pitch;
        L.pitch = g.pitch;
        L.padded = true;
        L.storage = pl->storage;
        L.scale = pl->scales.p;
        run_launch(pl, L, d_max_snr, d_best_v, static_cast<hipStream_t>(stream));
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_plan_tune(ftk_plan* pl, const float* d_frames, size_t frame_stride, float* d_max_snr,
                             int32_t* d_best_v, int reps, void* stream, int* tile_rows, int* queue, double* seconds,
                             char* err, size_t err_len) {
    try {
        if (!pl || !d_max_snr || !d_best_v) throw std::runtime_error("ftk_plan_tune: null argument");
        if (!d_frames && !pl->storage) throw std::runtime_error("ftk_plan_tune: no frames (give fp32 frames or store 16-bit ones)");
        check(hipSetDevice(pl->device), "hipSetDevice");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int storage = d_frames ? kStoreF32 : pl->storage;
        StackShape& target = storage == kStoreF32 ? pl->shape32 : pl->shape16;
        const StackShape keep = target;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        check(hipEventCreate(&e0), "create event");
        check(hipEventCreate(&e1), "create event");
        struct Events {
            hipEvent_t a, b;
            ~Events() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } held{e0, e1};
        StackShape best = keep;
        double best_t = 1e300;
        char inner[512] = {0};
        // the launch forms and tile heights the kernels take (a height the 16-bit staging
        // cannot hold for this plan is skipped)
        const StackShape cands[] = {{16, 0}, {8, 1}, {16, 1}, {32, 1}};
        for (const StackShape& c : cands) {
            if (!tile_rows_fit(pl->plan, storage, c.tile_rows)) continue;
            target = c;
            double t = 1e300;
            for (int r = 0; r <= std::max(1, reps); ++r) {  // the first run warms up
                check(hipEventRecord(e0, s), "record event");
                const int st = d_frames ? ftk_plan_run(pl, d_frames, frame_stride, d_max_snr, d_best_v, stream, inner, sizeof inner)
                                        : ftk_plan_run_stored(pl, d_max_snr, d_best_v, stream, inner, sizeof inner);
                if (st != FTK_OK) {
                    target = keep;
                    throw std::runtime_error(inner);
                }
                check(hipEventRecord(e1, s), "record event");
                check(hipEventSynchronize(e1), "stack pass");
                float ms = 0.0f;
                check(hipEventElapsedTime(&ms, e0, e1), "event time");
                if (r > 0) t = std::min(t, 1e-3 * ms);
            }
            if (t < best_t) {
                best_t = t;
                best = c;
            }
        }
        target = best;
        if (tile_rows) *tile_rows = best.tile_rows;
        if (queue) *queue = best.queue;
        if (seconds) *seconds = best_t;
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

namespace ftk {
// The 16-bit frames of a plan after ftk_plan_store: frame i element (x, y) at
// origin + i stride + y pitch + x, times scale[i] (device). False when none are stored.
bool plan_stored_frames(const ftk_plan* pl, const uint16_t** origin, long long* stride, int* pitch,
                        const float** scale) {
    if (!pl || !pl->storage || !pl->stored.p) return false;
    const PadGeometry& g = pl->pad16;
    *origin = pl->stored.p + static_cast<long long>(g.top) * g.pitch + g.left;
    *stride = static_cast<long long>(g.rows) * g.pitch;
    *pitch = g.pitch;
    *scale = pl->scales.p;
    return true;
}
}  // namespace ftk

extern "C" int ftk_plan_scales(const ftk_plan* pl, float* scales) {
    if (!pl || !scales || !pl->storage) return FTK_ERROR;
    if (hipSetDevice(pl->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return FTK_ERROR;
    return hipMemcpy(scales, pl->scales.p, sizeof(float) * pl->plan.n, hipMemcpyDeviceToHost) == hipSuccess ? FTK_OK
                                                                                                            : FTK_ERROR;
}

extern "C" int ftk_stack_fp64(int n, int w, int h, const float* d_frames, size_t frame_stride, int nv,
                              const int32_t* dx, const int32_t* dy, int min_count, int y0, int y1, double* d_max_snr,
                              int32_t* d_best_v, void* stream, char* err, size_t err_len) {
    try {
        if (!d_frames || !dx || !dy || !d_max_snr || !d_best_v) throw std::runtime_error("ftk_stack_fp64: null argument");
        if (n < 1 || w < 1 || h < 1 || nv < 1 || min_count < 1) throw std::runtime_error("ftk_stack_fp64: bad sizes");
        if (y0 < 0 || y1 > h || y1 < y0) throw std::runtime_error("ftk_stack_fp64: row range outside the frame");
        if (frame_stride < static_cast<size_t>(w) * h) throw std::runtime_error("ftk_stack_fp64: frame stride smaller than a frame");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t cnt = static_cast<size_t>(nv) * n;
        DeviceBuffer<int> ddx, ddy;
        ddx.reserve(cnt, "allocate shifts");
        ddy.reserve(cnt, "allocate shifts");
        check(hipMemcpyAsync(ddx.p, dx, cnt * sizeof(int), hipMemcpyHostToDevice, s), "upload shifts");
        check(hipMemcpyAsync(ddy.p, dy, cnt * sizeof(int), hipMemcpyHostToDevice, s), "upload shifts");
        check(launch_fp64(d_frames, static_cast<long long>(frame_stride), n, w, h, nv, ddx.p, ddy.p, min_count, y0, y1,
                          d_max_snr, d_best_v, s),
              "fp64 kernel");
        check(hipStreamSynchronize(s), "fp64 kernel");  // the tables are freed on return
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" double ftk_plan_units(const ftk_plan* pl) { return pl ? pl->plan.units : 0.0; }
extern "C" int ftk_plan_blocks(const ftk_plan* pl) { return pl ? pl->plan.nb : 0; }
extern "C" void ftk_plan_destroy(ftk_plan* pl) { delete pl; }

extern "C" int ftk_copy_f32(const float* d_src, float* d_dst, size_t count, void* stream) {
    return launch_copy(d_src, d_dst, static_cast<long long>(count), static_cast<hipStream_t>(stream)) == hipSuccess
               ? FTK_OK
               : FTK_ERROR;
}

// ------------------------------------------------------------------ host emulation
// The kernel's algorithm on the CPU: same plan, tiles, boxes, -0.0 padding of reads
// outside the frame, coverage counted from the full-frame bounds only on tiles that
// are not interior, and the packed-key maximum. Shares no code with the oracle.
extern "C" int ftk_plan_shape(int n, int w, int h, int nv, const int32_t* dx, const int32_t* dy, int min_count,
                              int* shape, char* err, size_t err_len) {
    try {
        if (!shape) throw std::runtime_error("ftk_plan_shape: null argument");
        const StackPlan p = make_stack_plan(n, w, h, nv, dx, dy, min_count);
        int lat_motions = 0;
        for (size_t q = 0; q < static_cast<size_t>(p.nb_lat) * kVel; ++q) lat_motions += p.motion[q] >= 0;
        shape[0] = p.nb;
        shape[1] = p.nb_lat;
        shape[2] = lat_motions;
        shape[3] = p.max_sy;
        shape[4] = p.nseg;
        shape[5] = p.blocks_per_seg;
        shape[6] = p.shadowed;
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}

extern "C" int ftk_emulate_pass(int n, int w, int h, const float* frames, int nv, const int32_t* dx, const int32_t* dy,
                                int min_count, float* max_snr, int32_t* best_v, char* err, size_t err_len) {
    try {
        StackPlan p = make_stack_plan(n, w, h, nv, dx, dy, min_count);
        const size_t plane = static_cast<size_t>(w) * h;
        std::vector<unsigned long long> keys(plane, 0ull);
        const int tiles = p.tiles_x * p.tiles_y;
        for (int t = 0; t < tiles; ++t) {
            const int x0 = (t % p.tiles_x) * kTileX, y0 = (t / p.tiles_x) * kTileY;
            std::vector<float> box(static_cast<size_t>(kLdsRows) * kPitch);
            std::vector<float> acc(static_cast<size_t>(kVel) * kTileX * kTileY);
            for (int seg = 0; seg < p.nseg; ++seg) {
                std::vector<float> best(static_cast<size_t>(kTileX) * kTileY, -std::numeric_limits<float>::infinity());
                std::vector<int> arg(static_cast<size_t>(kTileX) * kTileY, -1);
                const int b0 = seg * p.blocks_per_seg, b1 = std::min(p.nb, b0 + p.blocks_per_seg);
                for (int blk = b0; blk < b1; ++blk) {
                    const int* ub = &p.ubox[static_cast<size_t>(blk) * 4];
                    const bool interior = x0 + ub[0] >= 0 && x0 + kTileX - 1 + ub[2] < w && y0 + ub[1] >= 0 &&
                                          y0 + kTileY - 1 + ub[3] < h;
                    std::fill(acc.begin(), acc.end(), 0.0f);
                    for (int i = 0; i < n; ++i) {
                        const int* bx = &p.box[(static_cast<size_t>(blk) * n + i) * 4];
                        const int bx0 = x0 + bx[0], by0 = y0 + bx[1];
                        for (int r = 0; r < kTileY + bx[3]; ++r)
                            for (int c = 0; c < kTileX + bx[2]; ++c) {
                                const int gx = bx0 + c, gy = by0 + r;
                                box[static_cast<size_t>(r) * kPitch + c] =
                                    (gx >= 0 && gx < w && gy >= 0 && gy < h)
                                        ? frames[static_cast<size_t>(i) * plane + static_cast<size_t>(gy) * w + gx]
                                        : -0.0f;
                            }
                        const int* sh = &p.shift[(static_cast<size_t>(blk) * n + i) * 2 * kVel];
                        if (blk < p.nb_lat && interior) {
                            // the lattice kernel: column a of rows 0 .. 3 + 2 (extra pairs)
                            // read once per lane row group, rows c_b .. c_b + 3 added to slot
                            // 4 a + b (an unread row reads as NaN here, so a table that asked
                            // for one would change the result)
                            const uint32_t wx = static_cast<uint32_t>(p.lat[(static_cast<size_t>(blk) * n + i) * 2]);
                            const uint32_t wy = static_cast<uint32_t>(p.lat[(static_cast<size_t>(blk) * n + i) * 2 + 1]);
                            const int rows = 4 + 2 * static_cast<int>((wy >> 16) & 3u);
                            for (int a = 0; a < kLatSide; ++a)
                                for (int b = 0; b < kLatSide; ++b) {
                                    const int k = kLatSide * a + b;
                                    const int cb = static_cast<int>((wy >> (4 * b)) & 0xfu);
                                    const int cx = static_cast<int>((wx >> (8 * a)) & 0xffu);
                                    for (int ly = 0; ly < kTileY; ++ly)
                                        for (int lx = 0; lx < kTileX; ++lx) {
                                            const int q = ly % kRowsPerWave + cb;  // row of the lane's column read
                                            const float v = q < rows ? box[static_cast<size_t>(ly + cb) * kPitch + lx + cx]
                                                                     : std::numeric_limits<float>::quiet_NaN();
                                            acc[(static_cast<size_t>(k) * kTileY + ly) * kTileX + lx] += v;
                                        }
                                }
                            continue;
                        }
                        for (int k = 0; k < kVel; ++k)
                            for (int ly = 0; ly < kTileY; ++ly)
                                for (int lx = 0; lx < kTileX; ++lx)
                                    acc[(static_cast<size_t>(k) * kTileY + ly) * kTileX + lx] +=
                                        box[static_cast<size_t>(ly + sh[kVel + k] - bx[1]) * kPitch + lx + sh[k] - bx[0]];
                    }
                    for (int k = 0; k < kVel; ++k) {
                        const int j = p.motion[static_cast<size_t>(blk) * kVel + k];
                        if (j < 0) continue;
                        for (int ly = 0; ly < kTileY; ++ly)
                            for (int lx = 0; lx < kTileX; ++lx) {
                                int c = n;
                                if (!interior) {
                                    c = 0;
                                    for (int i = 0; i < n; ++i) {
                                        const int* sh = &p.shift[(static_cast<size_t>(blk) * n + i) * 2 * kVel];
                                        const int sx = x0 + lx + sh[k], sy = y0 + ly + sh[kVel + k];
                                        c += (sx >= 0 && sx < w && sy >= 0 && sy < h) ? 1 : 0;
                                    }
                                }
                                if (c < min_count) continue;
                                const float snr = acc[(static_cast<size_t>(k) * kTileY + ly) * kTileX + lx] * p.inv_sqrt[static_cast<size_t>(c)];
                                const size_t o = static_cast<size_t>(ly) * kTileX + lx;
                                if (snr > best[o] || (snr == best[o] && j < arg[o])) {
                                    best[o] = snr;
                                    arg[o] = j;
                                }
                            }
                    }
                }
                for (int ly = 0; ly < kTileY; ++ly)
                    for (int lx = 0; lx < kTileX; ++lx) {
                        const int x = x0 + lx, y = y0 + ly;
                        const size_t o = static_cast<size_t>(ly) * kTileX + lx;
                        if (arg[o] < 0 || x >= w || y >= h) continue;
                        const unsigned long long key = pack_key(best[o], arg[o]);
                        unsigned long long& dst = keys[static_cast<size_t>(y) * w + x];
                        dst = std::max(dst, key);
                    }
            }
        }
        for (size_t q = 0; q < plane; ++q) {
            if (keys[q] == 0ull) {
                max_snr[q] = 0.0f;
                best_v[q] = -1;
                continue;
            }
            const float v = from_ordered_bits(static_cast<uint32_t>(keys[q] >> 32));
            max_snr[q] = std::isfinite(v) ? v : 0.0f;
            best_v[q] = static_cast<int32_t>(~static_cast<uint32_t>(keys[q] & 0xffffffffull));
        }
        return FTK_OK;
    } catch (const std::exception& e) {
        set_err(err, err_len, e.what());
        return FTK_ERROR;
    }
}
