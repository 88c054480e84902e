// Serving loop in the library: resident stores, HBM-budget admission and
// continuous batching of apply requests, driven by two native threads.
//
// A request is (store, x, k, y): x order x k FP64 row-major in device
// memory, complete when submitted; y order x k, written when the request
// completes.  The dispatcher thread waits for a free slot (a stream with its
// pass buffers) and takes every queued request of the store at the head of
// the queue, up to max_rhs columns and kMaxMerge requests, as one pass:
//  - one request: qh_apply on its own x and y;
//  - several: a gather launch puts their columns side by side in the slot's
//    X, one apply of all the columns, a scatter launch hands each request its
//    columns of the slot's Y.
// The completion thread polls the passes in flight, and for each finished one
// records the latency of its requests (submission to observed completion)
// and marks them done (or, for the closed-loop load's sessions, submits the
// session's next request at once).
//
// A right-hand-side column's arithmetic depends only on the store, the
// precision and the pass's arithmetic path (vector ALUs up to 4 columns per
// pass, matrix cores above), not on the other columns of its pass; gather
// and scatter copy values unchanged.
//
// No counterpart in the reference (src/qhmat has no request loop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device_common.hpp"
#include "device_store.hpp"
#include "server.hpp"

namespace qhamd {
namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Columns of the requests merged into one pass: request j's k_j columns sit
// at columns [col_j, col_j + k_j) of the pass's order x K operands.
constexpr int kMaxMerge = 64;
struct MergeTable {
    const double* src[kMaxMerge];  // gather: the requests' x
    double* dst[kMaxMerge];        // scatter: the requests' y
    std::int32_t col[kMaxMerge + 1];
    std::int32_t count, K;
    std::uint64_t order;
};

// The request owning pass column c (col[j] <= c < col[j + 1]).
__device__ __forceinline__ int merge_owner(const MergeTable& t, std::int32_t c) {
    int lo = 0, hi = t.count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t.col[mid] <= c) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One thread per element of the pass's order x K operand, along the rows.
__global__ __launch_bounds__(256) void gather_columns(MergeTable t, double* __restrict__ X) {
    const std::uint64_t e = std::uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (e >= t.order * std::uint64_t(t.K)) return;
    const std::uint64_t i = e / std::uint64_t(t.K);
    const std::int32_t c = std::int32_t(e - i * std::uint64_t(t.K));
    const int j = merge_owner(t, c);
    const std::int32_t kj = t.col[j + 1] - t.col[j];
    X[e] = t.src[j][i * std::uint64_t(kj) + std::uint64_t(c - t.col[j])];
}
__global__ __launch_bounds__(256) void scatter_columns(MergeTable t, const double* __restrict__ Y) {
    const std::uint64_t e = std::uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (e >= t.order * std::uint64_t(t.K)) return;
    const std::uint64_t i = e / std::uint64_t(t.K);
    const std::int32_t c = std::int32_t(e - i * std::uint64_t(t.K));
    const int j = merge_owner(t, c);
    const std::int32_t kj = t.col[j + 1] - t.col[j];
    t.dst[j][i * std::uint64_t(kj) + std::uint64_t(c - t.col[j])] = Y[e];
}

// Fills n doubles with a deterministic pseudo-random sequence in [-1, 1)
// (the load generator's right-hand sides).
__global__ __launch_bounds__(256) void fill_operand(double* __restrict__ x, std::uint64_t n,
                                                    std::uint64_t seed) {
    const std::uint64_t e = std::uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (e >= n) return;
    std::uint64_t z = (e + 1) * 0x9E3779B97F4A7C15ull + seed;  // splitmix64
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[e] = double(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

unsigned blocks_for(std::uint64_t n) { return unsigned((n + 255) / 256); }

}  // namespace

struct Server::Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    double* X = nullptr;
    double* Y = nullptr;
    void* ws = nullptr;
    std::uint64_t x_doubles = 0, ws_bytes = 0;
    bool busy = false;      // taken by the dispatcher
    bool launched = false;  // its pass is on the stream (done recorded after it)
    std::vector<Request> reqs;
};

Server::Server(const ServerConfig& cfg) : cfg_(cfg) {
    require(cfg.slots >= 1 && cfg.slots <= 16, "server slots must be 1..16");
    require(cfg.max_rhs >= 1 && cfg.max_rhs <= 4096, "server max_rhs must be 1..4096");
    require(cfg.precision == 0 || cfg.precision == 1, "unknown precision");
    const int devices = device_count();
    if (devices == 0) throw std::runtime_error("no CUDA device is available (no HIP device found)");
    require(cfg.device >= 0 && cfg.device < devices, "device index out of range");
    DeviceScope scope(cfg.device);
    for (int i = 0; i < cfg.slots; ++i) {
        auto slot = std::make_unique<Slot>();
        hip_check(hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming), "hipEventCreate");
        slots_.push_back(std::move(slot));
    }
    running_ = true;
    dispatcher_ = std::thread([this] { dispatch_loop(); });
    completer_ = std::thread([this] { complete_loop(); });
}

Server::~Server() {
    {
        std::unique_lock<std::mutex> lock(mu_);
        running_ = false;
        cv_.notify_all();
    }
    if (dispatcher_.joinable()) dispatcher_.join();
    if (completer_.joinable()) completer_.join();
    DeviceScope scope(cfg_.device);
    for (auto& s : slots_) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(s->X);
        (void)hipFree(s->Y);
        (void)hipFree(s->ws);
        (void)hipEventDestroy(s->done);
        (void)hipStreamDestroy(s->stream);
    }
    for (auto& r : stores_)
        if (r.store) r.store->pins.fetch_sub(1);
}

// Waits (lock held) until no pass is in flight and the dispatcher is parked.
void Server::drain(std::unique_lock<std::mutex>& lock) {
    paused_ = true;
    cv_.notify_all();
    cv_.wait(lock, [this] { return inflight_ == 0 && !dispatching_; });
}

int Server::admit(const DeviceStore* store) {
    require(store != nullptr, "null store");
    require(store->device == cfg_.device, "the store lives on another device than the server");
    std::unique_lock<std::mutex> lock(mu_);
    drain(lock);
    struct Resume {
        Server* s;
        ~Resume() {
            s->paused_ = false;
            s->cv_.notify_all();
        }
    } resume{this};
    const std::uint64_t k = std::uint64_t(cfg_.max_rhs);
    const std::uint64_t x_doubles = std::max(store->order * k, std::uint64_t(1));
    const std::uint64_t ws = std::max(apply_workspace_bytes(*store, std::int64_t(k)), std::uint64_t(8));
    std::uint64_t grow = 0;
    for (auto& s : slots_) {
        if (x_doubles > s->x_doubles) grow += 2 * (x_doubles - s->x_doubles) * sizeof(double);
        if (ws > s->ws_bytes) grow += ws - s->ws_bytes;
    }
    const std::uint64_t size = store->stats.device_bytes;
    if (resident_bytes_ + buffer_bytes_ + grow + size > cfg_.hbm_budget)
        throw std::invalid_argument("the store and its pass buffers exceed the serving HBM budget");
    DeviceScope scope(cfg_.device);
    for (auto& s : slots_) {
        if (x_doubles > s->x_doubles) {
            hip_check(hipFree(s->X), "hipFree");
            hip_check(hipFree(s->Y), "hipFree");
            s->X = s->Y = nullptr;
            buffer_bytes_ -= 2 * s->x_doubles * sizeof(double);
            s->x_doubles = 0;
            hip_check(hipMalloc(&s->X, x_doubles * sizeof(double)), "hipMalloc");
            hip_check(hipMalloc(&s->Y, x_doubles * sizeof(double)), "hipMalloc");
            s->x_doubles = x_doubles;
            buffer_bytes_ += 2 * x_doubles * sizeof(double);
        }
        if (ws > s->ws_bytes) {
            hip_check(hipFree(s->ws), "hipFree");
            s->ws = nullptr;
            buffer_bytes_ -= s->ws_bytes;
            s->ws_bytes = 0;
            hip_check(hipMalloc(&s->ws, ws), "hipMalloc");
            s->ws_bytes = ws;
            buffer_bytes_ += ws;
        }
    }
    store->pins.fetch_add(1);
    stores_.push_back(Resident{store, size});
    resident_bytes_ += size;
    return int(stores_.size() - 1);
}

void Server::evict(int id) {
    std::unique_lock<std::mutex> lock(mu_);
    require(id >= 0 && std::size_t(id) < stores_.size() && stores_[id].store, "unknown store id");
    drain(lock);
    // queued requests of the store fail
    std::deque<Request> kept;
    for (auto& r : queue_) {
        if (r.store == id) finish_locked(r, false);
        else kept.push_back(r);
    }
    queue_.swap(kept);
    stores_[id].store->pins.fetch_sub(1);
    resident_bytes_ -= stores_[id].bytes;
    stores_[id] = Resident{};
    paused_ = false;
    cv_.notify_all();
}

std::uint64_t Server::submit(int id, const double* x, std::int64_t k, double* y) {
    std::unique_lock<std::mutex> lock(mu_);
    require(id >= 0 && std::size_t(id) < stores_.size() && stores_[id].store, "unknown store id");
    require(x != nullptr && y != nullptr, "null operand");
    require(k >= 1 && k <= cfg_.max_rhs, "request width must be 1..max_rhs");
    Request r{id, x, y, k, next_ticket_++, now_s(), -1};
    tickets_[r.ticket] = kPending;
    queue_.push_back(r);
    cv_.notify_all();
    return r.ticket;
}

int Server::wait(std::uint64_t ticket, double timeout_s) {
    std::unique_lock<std::mutex> lock(mu_);
    auto it = tickets_.find(ticket);
    require(it != tickets_.end(), "unknown ticket");
    const auto deadline = std::chrono::steady_clock::now() +
                          std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                              std::chrono::duration<double>(timeout_s));
    while (it->second == kPending) {
        if (timeout_s >= 0) {
            if (done_cv_.wait_until(lock, deadline) == std::cv_status::timeout &&
                it->second == kPending)
                return kPending;
        } else {
            done_cv_.wait(lock);
        }
        it = tickets_.find(ticket);
    }
    const int state = it->second;
    tickets_.erase(it);
    return state;
}

// Completion of one request (lock held): its ticket, or its session's next
// request.
void Server::finish_locked(const Request& r, bool ok) {
    if (r.session < 0) {
        tickets_[r.ticket] = ok ? kDone : kFailed;
        done_cv_.notify_all();
        return;
    }
    Session& s = sessions_[std::size_t(r.session)];
    if (!ok) load_errors_++;
    if (ok && s.left > 0) {
        s.left--;
        Request n{s.store, s.x, s.y, s.k, next_ticket_++, now_s(), r.session};
        queue_.push_back(n);
        cv_.notify_all();
    } else {
        if (--open_sessions_ == 0) done_cv_.notify_all();
    }
}

void Server::dispatch_loop() {
    DeviceScope scope(cfg_.device);
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
        Slot* slot = nullptr;
        cv_.wait(lock, [&] {
            if (!running_) return true;
            if (paused_ || queue_.empty()) return false;
            for (auto& s : slots_)
                if (!s->busy) return true;
            return false;
        });
        if (!running_) return;
        for (auto& s : slots_)
            if (!s->busy) {
                slot = s.get();
                break;
            }
        // the head request's store: every queued request of it that fits
        const int id = queue_.front().store;
        std::vector<Request> take;
        std::int64_t K = 0;
        std::deque<Request> kept;
        while (!queue_.empty()) {
            Request r = queue_.front();
            queue_.pop_front();
            if (r.store == id && K + r.k <= cfg_.max_rhs && int(take.size()) < kMaxMerge) {
                K += r.k;
                take.push_back(r);
            } else {
                kept.push_back(r);
            }
        }
        queue_.swap(kept);
        const DeviceStore* store = stores_[std::size_t(id)].store;
        if (!store) {  // evicted while a session still ran on it
            for (const auto& r : take) finish_locked(r, false);
            continue;
        }
        slot->busy = true;
        slot->reqs = take;
        dispatching_ = true;
        lock.unlock();
        bool ok = true;
        try {
            launch_pass(*slot, *store, take, K);
        } catch (const std::exception&) {
            ok = false;
        }
        lock.lock();
        dispatching_ = false;
        passes_++;
        if (ok) {
            slot->launched = true;
            inflight_++;
        } else {
            for (const auto& r : take) finish_locked(r, false);
            slot->busy = false;
            slot->reqs.clear();
        }
        cv_.notify_all();
    }
}

void Server::launch_pass(Slot& slot, const DeviceStore& store, const std::vector<Request>& take,
                         std::int64_t K) {
    const int prec = cfg_.precision;
    if (take.size() == 1) {
        const Request& r = take[0];
        apply(store, r.x, r.k, r.k, r.y, r.k, prec, slot.stream, slot.ws, slot.ws_bytes);
    } else {
        MergeTable t{};
        t.count = std::int32_t(take.size());
        t.K = std::int32_t(K);
        t.order = store.order;
        std::int32_t c = 0;
        for (std::size_t j = 0; j < take.size(); ++j) {
            t.src[j] = take[j].x;
            t.dst[j] = take[j].y;
            t.col[j] = c;
            c += std::int32_t(take[j].k);
        }
        t.col[take.size()] = c;
        const std::uint64_t n = store.order * std::uint64_t(K);
        hipLaunchKernelGGL(gather_columns, dim3(blocks_for(n)), dim3(256), 0, slot.stream, t, slot.X);
        hip_check(hipGetLastError(), "gather launch");
        apply(store, slot.X, K, K, slot.Y, K, prec, slot.stream, slot.ws, slot.ws_bytes);
        hipLaunchKernelGGL(scatter_columns, dim3(blocks_for(n)), dim3(256), 0, slot.stream, t,
                           static_cast<const double*>(slot.Y));
        hip_check(hipGetLastError(), "scatter launch");
    }
    hip_check(hipEventRecord(slot.done, slot.stream), "hipEventRecord");
}

void Server::complete_loop() {
    DeviceScope scope(cfg_.device);
    std::vector<Slot*> finished;
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return inflight_ > 0 || !running_; });
            if (!running_ && inflight_ == 0) return;
        }
        // poll the slots whose passes were launched
        finished.clear();
        for (auto& s : slots_) {
            bool launched;
            {
                std::lock_guard<std::mutex> g(mu_);
                launched = s->launched;
            }
            if (!launched) continue;
            const hipError_t e = hipEventQuery(s->done);
            if (e == hipErrorNotReady) continue;
            finished.push_back(s.get());
            if (e != hipSuccess) (void)hipGetLastError();
        }
        if (finished.empty()) {
            std::this_thread::yield();
            continue;
        }
        const double t = now_s();
        std::lock_guard<std::mutex> g(mu_);
        for (Slot* s : finished) {
            for (const auto& r : s->reqs) {
                latencies_ms_.push_back(float((t - r.t_submit) * 1e3));
                completed_++;
                completed_rhs_ += std::uint64_t(r.k);
                finish_locked(r, true);
            }
            s->reqs.clear();
            s->launched = false;
            s->busy = false;
            inflight_--;
        }
        cv_.notify_all();
    }
}

ServerStats Server::stats() {
    std::vector<float> lat;
    ServerStats out{};
    {
        std::lock_guard<std::mutex> g(mu_);
        lat = latencies_ms_;
        out.completed = completed_;
        out.completed_rhs = completed_rhs_;
        out.passes = passes_;
        out.resident_bytes = resident_bytes_;
        out.buffer_bytes = buffer_bytes_;
    }
    if (!lat.empty()) {
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {  // linear interpolation between order statistics
            const double at = p / 100.0 * double(lat.size() - 1);
            const std::size_t i = std::size_t(at);
            const double f = at - double(i);
            return i + 1 < lat.size() ? double(lat[i]) * (1 - f) + double(lat[i + 1]) * f : double(lat[i]);
        };
        out.p50_ms = pct(50);
        out.p99_ms = pct(99);
        double sum = 0;
        for (float v : lat) sum += v;
        out.mean_ms = sum / double(lat.size());
    }
    return out;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> g(mu_);
    latencies_ms_.clear();
    completed_ = completed_rhs_ = passes_ = 0;
}

LoadResult Server::closed_loop(const std::vector<int>& ids, const std::vector<std::int64_t>& ks,
                               double rate, double duration, int session, std::uint64_t seed) {
    require(!ids.empty() && !ks.empty(), "the load needs stores and widths");
    require(rate > 0 && duration > 0 && session >= 1, "the load needs a rate, a duration and a session length");
    for (std::int64_t k : ks) require(k >= 1 && k <= cfg_.max_rhs, "load width must be 1..max_rhs");
    DeviceScope scope(cfg_.device);
    // operand and result buffers per (store, width); sessions on one pair
    // share them (the results are not read)
    std::vector<double*> xs, ys;
    auto free_all = [&] {
        for (double* p : xs) (void)hipFree(p);
        for (double* p : ys) (void)hipFree(p);
    };
    std::vector<std::uint64_t> orders;
    {
        std::lock_guard<std::mutex> g(mu_);
        for (int id : ids) {
            require(id >= 0 && std::size_t(id) < stores_.size() && stores_[id].store, "unknown store id");
            orders.push_back(stores_[id].store->order);
        }
    }
    try {
        for (std::size_t a = 0; a < ids.size(); ++a)
            for (std::size_t b = 0; b < ks.size(); ++b) {
                const std::uint64_t n = orders[a] * std::uint64_t(ks[b]);
                double *x = nullptr, *y = nullptr;
                hip_check(hipMalloc(&x, n * sizeof(double)), "hipMalloc");
                xs.push_back(x);
                hip_check(hipMalloc(&y, n * sizeof(double)), "hipMalloc");
                ys.push_back(y);
                hipLaunchKernelGGL(fill_operand, dim3(blocks_for(n)), dim3(256), 0, nullptr, x, n,
                                   seed * 1315423911ull + xs.size());
                hip_check(hipGetLastError(), "fill launch");
            }
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        // one request per (store, width) first: first launches of each shape,
        // dropped from the statistics
        std::vector<std::uint64_t> warm;
        for (std::size_t a = 0; a < ids.size(); ++a)
            for (std::size_t b = 0; b < ks.size(); ++b)
                warm.push_back(submit(ids[a], xs[a * ks.size() + b], ks[b], ys[a * ks.size() + b]));
        for (auto t : warm)
            if (wait(t, -1.0) != kDone) throw std::runtime_error("a warm-up request failed");
        reset_stats();

        // session starts at Poisson times, rate / session per second
        std::mt19937_64 rng(seed);
        std::exponential_distribution<double> gap(rate / double(session));
        std::vector<double> starts;
        for (double t = gap(rng); t < duration; t += gap(rng)) starts.push_back(t);
        std::vector<Session> plan;
        for (std::size_t i = 0; i < starts.size(); ++i) {
            const std::size_t a = std::size_t(rng() % ids.size()), b = std::size_t(rng() % ks.size());
            plan.push_back(Session{ids[a], ks[b], xs[a * ks.size() + b], ys[a * ks.size() + b], session});
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            sessions_ = plan;
            open_sessions_ = int(plan.size());
            load_errors_ = 0;
        }
        const double t0 = now_s();
        for (std::size_t i = 0; i < plan.size(); ++i) {
            const double delay = starts[i] - (now_s() - t0);
            if (delay > 0) std::this_thread::sleep_for(std::chrono::duration<double>(delay));
            std::lock_guard<std::mutex> g(mu_);
            Session& s = sessions_[i];
            s.left--;
            queue_.push_back(Request{s.store, s.x, s.y, s.k, next_ticket_++, now_s(), int(i)});
            cv_.notify_all();
        }
        {
            std::unique_lock<std::mutex> lock(mu_);
            done_cv_.wait(lock, [&] { return open_sessions_ == 0; });
        }
        const double elapsed = now_s() - t0;
        LoadResult out{};
        const ServerStats st = stats();
        out.offered_rate = rate;
        out.offered_rate_drawn = double(plan.size()) * double(session) / duration;
        out.sessions = plan.size();
        out.requests = st.completed;
        out.seconds = elapsed;
        out.throughput_rps = double(st.completed) / elapsed;
        out.passes = st.passes;
        out.p50_ms = st.p50_ms;
        out.p99_ms = st.p99_ms;
        out.mean_ms = st.mean_ms;
        {
            std::lock_guard<std::mutex> g(mu_);
            out.errors = load_errors_;
            sessions_.clear();
        }
        free_all();
        return out;
    } catch (...) {
        // the sessions' requests may still be in flight: let them finish
        {
            std::unique_lock<std::mutex> lock(mu_);
            done_cv_.wait(lock, [&] { return open_sessions_ == 0; });
            sessions_.clear();
        }
        free_all();
        throw;
    }
}

}  // namespace qhamd
