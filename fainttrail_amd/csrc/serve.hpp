// Device-resident search server: a continuous-batching request loop over
// DeviceSearch (search_device.hpp). No reference counterpart (the reference runs one
// search per process, src/search.cpp:716-809).
//
// Jobs (frames already in HBM, times, search parameters) queue in arrival order. A
// worker slot that is free admits the first queued job whose HBM reservation fits the
// memory budget next to the jobs already running and the idle searchers kept for
// reuse (idle searchers are evicted, oldest first, to make room). A smaller job may
// thus overtake a larger one that does not fit yet, but only a bounded number of
// times: once the head of the queue has been overtaken kMaxOvertake times, nothing
// else is admitted before it, so no job starves. A job reserves its footprint
// estimate (search_footprint), or the bytes a reused searcher of its shape already
// holds if that is more; when the job ends the searcher goes back to the idle list at
// the bytes it actually holds (DeviceSearch::device_bytes). Each slot runs its job's
// batched search on its own streams, so the stage builds, stacking launches and peak
// transfers of different jobs overlap on the device, and jobs of the same frame size
// whose stacking launches meet are stacked in one persistent launch (fuse.hpp). A job's
// result depends on its inputs only: the same alone and in any mix.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fuse.hpp"
#include "search_device.hpp"

namespace ftk {

// Device bytes of a DeviceSearch over n frames of w x h and of its batched run with
// nv motions in groups of G passes (two sets of pass buffers when the run has more
// than one group), for the given frame storage, with a candidate stage for up to
// max_candidates candidates (0: none): what admission reserves.
double search_footprint(int n, int w, int h, long long nv, int G, int passes, int storage, int max_candidates);

struct ServerJob {
    long long id = 0;
    int n = 0, w = 0, h = 0;
    const float* raw = nullptr;
    long long stride = 0;
    std::vector<double> mjd;
    SearchConfig cfg;
    long long nv = 0;
    int passes = 1;
    double bytes = 0.0;                  // footprint estimate (search_footprint)
    double reserved = 0.0;               // what admission reserved: max(estimate, reused searcher's bytes)
    double held = 0.0;                   // device bytes the job's searcher held when it ended
    int overtaken = 0;                   // later jobs admitted while this one waited at the head
    double t_submit = 0, t_admit = 0, t_end = 0;  // seconds on the server's clock
    int status = 0;                      // 0 queued, 1 running, 2 done, 3 failed
    std::string error;
    SearchOutput out;
};

class SearchServer {
public:
    SearchServer(int device, double budget_bytes, int slots);
    ~SearchServer();  // finishes the running jobs, fails the queued ones
    SearchServer(const SearchServer&) = delete;
    SearchServer& operator=(const SearchServer&) = delete;

    long long submit(const float* raw, long long stride, int n, int w, int h, const double* mjd, const SearchConfig& cfg);
    // Waits up to timeout_s (< 0: forever) for the job to end; false on time-out.
    bool wait(long long id, double timeout_s);
    // The job (nullptr if unknown); valid until release(id).
    const ServerJob* job(long long id);
    void release(long long id);
    struct Stats {
        double used = 0, peak = 0, budget = 0, idle_bytes = 0;
        int running = 0, queued = 0, completed = 0, failed = 0, evicted = 0, max_running = 0;
        long long fused_launches = 0, fused_passes = 0, stack_launches = 0;
    };
    // Cross-job fusion of persistent stacking launches (on by default, 3 ms window).
    void set_fusion(bool on, double window_s);
    Stats stats();
    double now() const;

private:
    struct Idle {
        int n, w, h;
        std::unique_ptr<DeviceSearch> ds;
        double bytes;
    };
    void worker();
    // With mu_ held: the queue position of the job to admit next, or -1.
    int pick_admissible();
    static constexpr int kMaxOvertake = 16;

    int device_;
    double budget_;
    int slots_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<ServerJob*> queue_;
    std::map<long long, std::unique_ptr<ServerJob>> jobs_;
    std::vector<Idle> idle_;
    double used_ = 0.0, peak_ = 0.0;
    int running_ = 0, max_running_ = 0, completed_ = 0, failed_ = 0, evicted_ = 0;
    long long next_id_ = 1;
    bool stop_ = false;
    double t0_ = 0.0;
    StackFuser fuser_{0.003};
    bool fuse_ = true;
    std::vector<std::thread> workers_;
};

}  // namespace ftk
