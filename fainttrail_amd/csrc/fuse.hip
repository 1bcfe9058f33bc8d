// Cross-job stacking fusion for the search server (see fuse.hpp).
#include "fuse.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>

namespace ftk {

namespace {
constexpr size_t kSlots = 4;
void fcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

bool StackFuser::Key::operator==(const Key& o) const {
    if (storage != o.storage || pre != o.pre || padded != o.padded || tile_rows != o.tile_rows) return false;
    for (int k = 0; k < kArgsKinds; ++k)
        if (ntiles[k] != o.ntiles[k] || nseg[k] != o.nseg[k]) return false;
    return true;
}

StackFuser::Key StackFuser::key_of(const StackArgs* host, int stride) {
    Key k{};
    const StackArgs& r = host[kArgsRing * stride];
    k.storage = r.storage;
    k.pre = r.pre;
    k.padded = r.padded;
    k.tile_rows = r.tile_rows;
    for (int kind = 0; kind < kArgsKinds; ++kind) {
        const StackArgs& a = host[kind * stride];
        // the work items of the kind as the persistent launch counts them
        k.ntiles[kind] = a.nseg > 0 ? a.ntiles : 0;
        k.nseg[kind] = k.ntiles[kind] > 0 ? a.nseg : 0;
    }
    return k;
}

StackFuser::StackFuser(double window_s) : window_s_(std::max(0.0, window_s)), slots_(kSlots) {}

StackFuser::~StackFuser() {
    for (Slot& sl : slots_) {
        if (sl.used && sl.done) (void)hipEventSynchronize(sl.done);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.dev) (void)hipFree(sl.dev);
        if (sl.counters) (void)hipFree(sl.counters);
        if (sl.host) (void)hipHostFree(sl.host);
    }
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
}

void StackFuser::set_window(double window_s) {
    std::lock_guard<std::mutex> lk(mu_);
    window_s_ = std::max(0.0, window_s);
}

void StackFuser::enter(int w, int h) {
    std::lock_guard<std::mutex> lk(mu_);
    active_.emplace_back(w, h);
}

void StackFuser::leave(int w, int h) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = std::find(active_.begin(), active_.end(), std::make_pair(w, h));
        if (it != active_.end()) active_.erase(it);
    }
    cv_.notify_all();  // a submitter waiting for this job launches without it
}

StackFuser::Stats StackFuser::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

hipEvent_t StackFuser::take_event() {
    if (!events_.empty()) {
        hipEvent_t e = events_.back();
        events_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    fcheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "create event");
    return e;
}

void StackFuser::give_event(hipEvent_t e) { events_.push_back(e); }

void StackFuser::stack(const StackArgs* host, const StackArgs* dev, int count, int stride, int* counters, hipStream_t s,
                       int w, int h) {
    Sub me{host, dev, count, stride, counters, s, nullptr, key_of(host, stride), w, h};
    std::unique_lock<std::mutex> lk(mu_);
    me.ready = take_event();
    // the job's stage builds and earlier work are ordered on s before this event
    const hipError_t rec = hipEventRecord(me.ready, s);
    if (rec != hipSuccess) {
        give_event(me.ready);
        fcheck(rec, "record event");
    }
    pending_.push_back(&me);
    cv_.notify_all();
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(window_s_);
    while (!me.done) {
        if (me.taken) {  // another submitter is launching this one
            cv_.wait(lk);
            continue;
        }
        // jobs of this frame size that are running and have not submitted may still join
        int running = 0, waiting = 0;
        for (const auto& a : active_) running += (a.first == w && a.second == h) ? 1 : 0;
        for (const Sub* p : pending_) waiting += (!p->taken && p->w == w && p->h == h) ? 1 : 0;
        if (running > waiting && std::chrono::steady_clock::now() < deadline) {
            cv_.wait_until(lk, deadline);
            continue;
        }
        // launch every pending submission of this shape, this one last (its stream runs it)
        std::vector<Sub*> mine;
        for (Sub* p : pending_)
            if (p != &me && !p->taken && p->key == me.key) {
                p->taken = true;
                mine.push_back(p);
            }
        me.taken = true;
        mine.push_back(&me);
        pending_.erase(std::remove_if(pending_.begin(), pending_.end(), [](const Sub* p) { return p->taken; }),
                       pending_.end());
        lk.unlock();
        hipError_t err = hipSuccess;
        try {
            err = launch(mine);
        } catch (const std::exception&) {
            err = hipErrorUnknown;
        }
        lk.lock();
        for (Sub* p : mine) {
            p->err = err;
            p->done = true;
        }
        if (err == hipSuccess) {
            ++stats_.launches;
            if (mine.size() > 1) {
                ++stats_.fused_launches;
                for (const Sub* p : mine) stats_.fused_passes += p->count;
            }
        }
        cv_.notify_all();
    }
    give_event(me.ready);
    const hipError_t err = me.err;
    lk.unlock();
    fcheck(err, "stacking launch");
}

// subs.back() is the launching thread's own submission: the launch goes on its stream.
hipError_t StackFuser::launch(std::vector<Sub*>& subs) {
    Sub& L = *subs.back();
    if (subs.size() == 1) return launch_stack_queue(L.host, L.dev, L.count, L.stride, L.counters, L.s);
    std::lock_guard<std::mutex> g(launch_mu_);
    int total = 0;
    for (const Sub* p : subs) total += p->count;
    Slot& sl = slots_[next_slot_++ % kSlots];
    // the slot's previous launch has read its descriptors and counters
    if (sl.used) fcheck(hipEventSynchronize(sl.done), "fused launch");
    const size_t need = static_cast<size_t>(kArgsKinds) * total;
    if (need > sl.cap) {
        if (sl.dev) fcheck(hipFree(sl.dev), "free descriptors");
        if (sl.host) fcheck(hipHostFree(sl.host), "free descriptors");
        sl.dev = nullptr;
        sl.host = nullptr;
        sl.cap = 0;
        const size_t cap = need + need / 2;
        fcheck(hipMalloc(reinterpret_cast<void**>(&sl.dev), cap * sizeof(StackArgs)), "allocate descriptors");
        fcheck(hipHostMalloc(reinterpret_cast<void**>(&sl.host), cap * sizeof(StackArgs), hipHostMallocDefault),
               "allocate descriptors");
        sl.cap = cap;
    }
    if (!sl.counters) fcheck(hipMalloc(reinterpret_cast<void**>(&sl.counters), 64 * sizeof(int)), "allocate counters");
    if (!sl.done) fcheck(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "create event");
    // descriptors of every pass, kind-major, `total` apart
    int off = 0;
    for (const Sub* p : subs) {
        for (int kind = 0; kind < kArgsKinds; ++kind)
            std::memcpy(sl.host + static_cast<size_t>(kind) * total + off, p->host + static_cast<size_t>(kind) * p->stride,
                        sizeof(StackArgs) * p->count);
        off += p->count;
    }
    // the launch waits for every job's builds, and every job's stream for the launch
    for (const Sub* p : subs)
        if (p != &L) fcheck(hipStreamWaitEvent(L.s, p->ready, 0), "stream wait");
    fcheck(hipMemcpyAsync(sl.dev, sl.host, need * sizeof(StackArgs), hipMemcpyHostToDevice, L.s), "upload descriptors");
    const hipError_t e = launch_stack_queue(sl.host, sl.dev, total, total, sl.counters, L.s);
    if (e != hipSuccess) return e;
    fcheck(hipEventRecord(sl.done, L.s), "record event");
    sl.used = true;
    for (const Sub* p : subs)
        if (p != &L) fcheck(hipStreamWaitEvent(p->s, sl.done, 0), "stream wait");
    return hipSuccess;
}

}  // namespace ftk
