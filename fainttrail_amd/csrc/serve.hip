// Device-resident search server (see serve.hpp).
#include "serve.hpp"
#include "refine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

#include "stack_plan.hpp"

namespace ftk {

namespace {
double clock_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

double search_footprint(int n, int w, int h, long long nv, int G, int passes, int storage, int max_candidates) {
    const double P = static_cast<double>(w) * h;
    const double pk_cap = std::max(65536.0, P / 8.0);
    // DeviceSearch: background-subtracted and filtered frames, four maps and the static
    // sky, peak buffer, null-screen tables; DeviceStage: residuals, flags, sky / scatter /
    // mask planes, small tables
    double base = 8.0 * n * P + 20.0 * P + 16.0 * pk_cap + 5.0 * n * P + 12.0 * P + 1048576.0 + 65536.0 * n +
                  16.0 * 3000.0 * passes;
    // small frames get more build lanes: stage scratch and a build target each
    base += (build_lanes(n, static_cast<long long>(w) * h, passes) - 1) * (9.0 * n * P + 12.0 * P + 65536.0 * n);
    // one pass of a group: frames (a padded layout is at most twice a frame for fp32;
    // 16-bit frames take half of that), keys and maps, plan rows (shift, box, LDS
    // offsets and lattice offsets: at most 54 ints per block and frame, blocks of 16
    // motions, some split or partly filled), peaks
    const double frame = storage == 0 ? 8.0 * P : 4.0 * P;
    const double tab = 54.0 * 4.0 * std::ceil(1.25 * static_cast<double>(nv) / kVel) * n;
    const double per_pass = n * frame + 24.0 * P + tab + 16.0 * pk_cap + 12.0 * n + 1024.0;
    // the real pass goes alone first, so any search with null runs uses both sets
    const double sets = (passes > 1 ? 2.0 : 1.0) * std::max(1, G) * per_pass;
    // per-run tables: the real rows, coverage tables (at most 2^26 entries each), orders
    const double run = tab + 2.0 * 4.0 * std::min(67108864.0, 16.0 * std::ceil(1.25 * nv / kVel) * (w + h)) +
                       4.0 * n * passes;
    // candidate stage (max_candidates > 0): the real pass's frames, residuals and mask,
    // the refinement's tables and figures (a grown buffer: 1.25 x, at least 64 KiB)
    const double mc = std::max(0, max_candidates);
    const double cand = mc > 0 ? 8.0 * n * P + 4.0 * P +
                                     std::max(65536.0, 1.25 * (8.0 * (4.0 * n + 4096.0 + 4.0 * mc + kRefOut * mc) +
                                                               4.0 * mc * n))
                               : 0.0;
    return base + sets + run + cand;
}

SearchServer::SearchServer(int device, double budget_bytes, int slots)
    : device_(device), budget_(budget_bytes), slots_(std::max(1, slots)) {
    if (hipSetDevice(device) != hipSuccess) throw std::runtime_error("HIP device " + std::to_string(device) + " not present");
    if (!(budget_ > 0.0)) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) throw std::runtime_error("HIP: memory info");
        budget_ = 0.85 * static_cast<double>(free_b);
    }
    t0_ = clock_s();
    for (int k = 0; k < slots_; ++k) workers_.emplace_back([this] { worker(); });
}

SearchServer::~SearchServer() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        for (ServerJob* j : queue_) {
            j->status = 3;
            j->error = "server stopped before the job was admitted";
            j->t_end = now();
            ++failed_;
        }
        queue_.clear();
    }
    cv_.notify_all();
    done_cv_.notify_all();
    for (auto& t : workers_) t.join();
    (void)hipSetDevice(device_);
    idle_.clear();
}

double SearchServer::now() const { return clock_s() - t0_; }

long long SearchServer::submit(const float* raw, long long stride, int n, int w, int h, const double* mjd,
                               const SearchConfig& cfg_in) {
    if (!raw || !mjd) throw std::runtime_error("server: null frames or times");
    if (n < 5 || w < 16 || h < 16) throw std::runtime_error("server: frames are too few or too small");
    if (stride < static_cast<long long>(w) * h) throw std::runtime_error("server: frame stride smaller than a frame");
    auto j = std::make_unique<ServerJob>();
    j->n = n;
    j->w = w;
    j->h = h;
    j->raw = raw;
    j->stride = stride;
    j->mjd.assign(mjd, mjd + n);
    for (int i = 1; i < n; ++i)
        if (j->mjd[static_cast<size_t>(i)] < j->mjd[static_cast<size_t>(i) - 1])
            throw std::runtime_error("server: frames must be sorted by time");
    j->cfg = cfg_in;
    // the persistent stacking launch (fusable with other jobs') unless the job asks
    if (j->cfg.queue < 0) j->cfg.queue = 1;
    const double span = (j->mjd.back() - j->mjd.front()) * 24.0;
    double dv = 0.0;
    // motions: every (inverse distance, velocity) pair
    j->nv = static_cast<long long>(velocity_grid(span, j->cfg.vmin_pxh, j->cfg.vmax_pxh, j->cfg.grid_tol_px, &dv).size()) *
            static_cast<long long>(inverse_distance_grid(j->cfg, n, nullptr).size());
    j->passes = 1 + (j->cfg.threshold_snr > 0.0 ? 0 : std::max(0, j->cfg.null_runs));
    // group size: as asked, else up to 8 passes per launch, fewer if the budget needs it
    int G = j->cfg.passes_per_launch > 0 ? j->cfg.passes_per_launch : std::min(8, j->passes);
    G = std::max(1, std::min(G, j->passes));
    const int mc = j->cfg.refines() ? j->cfg.max_candidates : 0;
    auto footprint = [&](int g) { return search_footprint(n, w, h, j->nv, g, j->passes, j->cfg.storage, mc); };
    while (G > 1 && footprint(G) > budget_) --G;
    j->cfg.passes_per_launch = G;
    j->bytes = footprint(G);
    if (j->bytes > budget_)
        throw std::runtime_error("server: the job needs " + std::to_string(j->bytes / 1e9) + " GB, more than the " +
                                 std::to_string(budget_ / 1e9) + " GB budget");
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("server: stopped");
    j->id = next_id_++;
    j->t_submit = now();
    ServerJob* p = j.get();
    jobs_[p->id] = std::move(j);
    queue_.push_back(p);
    cv_.notify_all();
    return p->id;
}

int SearchServer::pick_admissible() {
    if (queue_.empty() || running_ >= slots_) return -1;
    double idle_total = 0.0;
    for (const Idle& d : idle_) idle_total += d.bytes;
    for (size_t q = 0; q < queue_.size(); ++q) {
        const ServerJob* j = queue_[q];
        // a reused searcher of the job's shape keeps what it holds; idle searchers can be
        // evicted (or reused) to make room
        double need = j->bytes;
        for (const Idle& d : idle_)
            if (d.n == j->n && d.w == j->w && d.h == j->h) {
                need = std::max(need, d.bytes);
                break;
            }
        if (used_ - idle_total + need <= budget_) return static_cast<int>(q);
        // the head has waited through enough overtaking: it goes next
        if (q == 0 && j->overtaken >= kMaxOvertake) return -1;
    }
    return -1;
}

void SearchServer::worker() {
    (void)hipSetDevice(device_);
    hipStream_t stream = nullptr;
    (void)hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        int pos = -1;
        cv_.wait(lk, [&] { return stop_ || (pos = pick_admissible()) >= 0; });
        if (stop_) break;
        ServerJob* j = queue_[static_cast<size_t>(pos)];
        for (int q = 0; q < pos; ++q) ++queue_[static_cast<size_t>(q)]->overtaken;
        queue_.erase(queue_.begin() + pos);
        std::unique_ptr<DeviceSearch> ds;
        double held = 0.0;
        for (auto it = idle_.begin(); it != idle_.end(); ++it)
            if (it->n == j->n && it->w == j->w && it->h == j->h) {
                ds = std::move(it->ds);
                held = it->bytes;
                idle_.erase(it);
                break;
            }
        j->reserved = std::max(j->bytes, held);
        std::vector<std::unique_ptr<DeviceSearch>> evict;
        used_ -= held;
        while (used_ + j->reserved > budget_ && !idle_.empty()) {
            used_ -= idle_.front().bytes;
            evict.push_back(std::move(idle_.front().ds));
            idle_.erase(idle_.begin());
            ++evicted_;
        }
        used_ += j->reserved;
        peak_ = std::max(peak_, used_);
        j->status = 1;
        j->t_admit = now();
        ++running_;
        max_running_ = std::max(max_running_, running_);
        lk.unlock();
        evict.clear();  // device memory of evicted searchers is freed outside the lock
        int status = 2;
        std::string error;
        SearchOutput out;
        double held_now = 0.0;
        bool fuse = false;
        {
            std::lock_guard<std::mutex> g(mu_);
            fuse = fuse_;
        }
        if (fuse) fuser_.enter(j->w, j->h);
        try {
            if (!ds) ds = std::make_unique<DeviceSearch>(j->n, j->w, j->h);
            ds->set_fuser(fuse ? &fuser_ : nullptr);
            out = ds->run(j->raw, j->stride, j->mjd.data(), j->cfg, nullptr, nullptr, stream);
            held_now = ds->device_bytes();
        } catch (const std::exception& e) {
            status = 3;
            error = e.what();
            ds.reset();  // a searcher that failed is not reused
            (void)hipStreamSynchronize(stream);
        }
        if (fuse) fuser_.leave(j->w, j->h);
        lk.lock();
        j->out = std::move(out);
        j->error = error;
        j->t_end = now();
        j->held = held_now;
        j->status = status;
        --running_;
        if (status == 2) ++completed_;
        else ++failed_;
        used_ -= j->reserved;
        if (ds) {
            // kept for reuse at the bytes it holds, which may exceed this job's estimate
            // (its buffers only grow): later admissions count them
            idle_.push_back(Idle{j->n, j->w, j->h, std::move(ds), held_now});
            used_ += held_now;
            peak_ = std::max(peak_, used_);
        }
        done_cv_.notify_all();
        cv_.notify_all();
    }
    lk.unlock();
    if (stream) (void)hipStreamDestroy(stream);
}

bool SearchServer::wait(long long id, double timeout_s) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it == jobs_.end()) throw std::runtime_error("server: unknown job");
    ServerJob* j = it->second.get();
    auto done = [&] { return j->status >= 2; };
    if (timeout_s < 0) {
        done_cv_.wait(lk, done);
        return true;
    }
    return done_cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
}

const ServerJob* SearchServer::job(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    return it == jobs_.end() ? nullptr : it->second.get();
}

void SearchServer::release(long long id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = jobs_.find(id);
    if (it != jobs_.end() && it->second->status >= 2) jobs_.erase(it);
}

SearchServer::Stats SearchServer::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s;
    s.used = used_;
    s.peak = peak_;
    s.budget = budget_;
    for (const Idle& d : idle_) s.idle_bytes += d.bytes;
    s.running = running_;
    s.queued = static_cast<int>(queue_.size());
    s.completed = completed_;
    s.failed = failed_;
    s.evicted = evicted_;
    s.max_running = max_running_;
    const StackFuser::Stats f = fuser_.stats();
    s.fused_launches = f.fused_launches;
    s.fused_passes = f.fused_passes;
    s.stack_launches = f.launches;
    return s;
}

void SearchServer::set_fusion(bool on, double window_s) {
    std::lock_guard<std::mutex> lk(mu_);
    fuse_ = on;
    fuser_.set_window(window_s);
}

}  // namespace ftk
