// Cross-job stacking fusion for the search server: passes of different jobs whose
// persistent stacking launches have the same work shape (the same tiles and velocity
// segments per kind, frame storage, 16-bit rows per wavefront, padding and tile height)
// go into ONE persistent launch when their jobs reach their stacking at the same time.
// No reference counterpart (the reference runs one search per process).
//
// A job's worker thread hands its group's launch to stack() instead of launching it:
// the call records an event on the job's stream (its stage builds are done there) and
// waits, for at most the window, while another job of the same frame size is running
// and has not submitted yet. Then one of the waiting threads launches every pending
// submission of its shape together on its own stream (after the others' events), and
// the other jobs' streams wait for that launch. Each pass keeps its own descriptor,
// tables, frames and keys, and the kernel stacks every work item exactly as in a launch
// of its own, so a job's result is the same fused or alone.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

#include "stack_launch.hpp"

namespace ftk {

class StackFuser {
public:
    explicit StackFuser(double window_s);
    ~StackFuser();
    StackFuser(const StackFuser&) = delete;
    StackFuser& operator=(const StackFuser&) = delete;

    // A job over frames of w x h starts / ends (the jobs a submitter may wait for).
    void enter(int w, int h);
    void leave(int w, int h);
    // The persistent stacking launch of `count` passes (descriptors kind-major, `stride`
    // apart: host copies for the shapes, device copies for the kernel), ordered on s:
    // fused with other jobs' submissions of the same shape, or alone. Returns once the
    // launch is enqueued and s waits for it. Throws on a launch error.
    void stack(const StackArgs* host, const StackArgs* dev, int count, int stride, int* counters, hipStream_t s, int w,
               int h);

    void set_window(double window_s);
    struct Stats {
        long long launches = 0;        // persistent launches made through the fuser
        long long fused_launches = 0;  // of those, launches holding passes of two or more jobs
        long long fused_passes = 0;    // passes stacked in fused launches
    };
    Stats stats();

private:
    struct Key {
        int storage, pre, padded, tile_rows;
        int ntiles[kArgsKinds], nseg[kArgsKinds];
        bool operator==(const Key& o) const;
    };
    struct Sub {
        const StackArgs* host;
        const StackArgs* dev;
        int count, stride;
        int* counters;
        hipStream_t s;
        hipEvent_t ready;
        Key key;
        int w, h;
        bool taken = false, done = false;
        hipError_t err = hipSuccess;
    };
    // A launch slot: pinned staging and device copy of a fused launch's descriptors,
    // its queue counters, and the event that marks the launch (its end frees the slot).
    struct Slot {
        StackArgs* host = nullptr;
        StackArgs* dev = nullptr;
        size_t cap = 0;
        int* counters = nullptr;
        hipEvent_t done = nullptr;
        bool used = false;
    };
    static Key key_of(const StackArgs* host, int stride);
    hipError_t launch(std::vector<Sub*>& subs);  // without mu_
    hipEvent_t take_event();                     // with mu_ held
    void give_event(hipEvent_t e);               // with mu_ held

    std::mutex mu_;
    std::condition_variable cv_;
    double window_s_;
    std::vector<Sub*> pending_;
    std::vector<std::pair<int, int>> active_;  // frame sizes of the running jobs
    std::vector<hipEvent_t> events_;           // free events
    std::vector<Slot> slots_;
    size_t next_slot_ = 0;
    std::mutex launch_mu_;  // one fused launch enqueued at a time (slots, staging)
    Stats stats_;
};

}  // namespace ftk
