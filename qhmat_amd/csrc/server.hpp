// Serving loop of the library (server.hip): resident stores, HBM-budget
// admission, continuous batching of apply requests on native threads.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device_store.hpp"

namespace qhamd {

struct ServerConfig {
    int device = 0;
    int precision = 0;           // qh_precision
    int slots = 3;               // streams with their own pass buffers
    std::int64_t max_rhs = 64;   // columns per pass (and per request)
    std::uint64_t hbm_budget = 0;  // stores plus pass buffers
};

struct ServerStats {
    std::uint64_t completed, completed_rhs, passes, resident_bytes, buffer_bytes;
    double p50_ms, p99_ms, mean_ms;  // since the last reset; 0 when none
};

struct LoadResult {
    double offered_rate, offered_rate_drawn;
    std::uint64_t sessions, requests, passes, errors;
    double seconds, throughput_rps, p50_ms, p99_ms, mean_ms;
};

class Server {
  public:
    static constexpr int kPending = 0, kDone = 1, kFailed = 2;

    explicit Server(const ServerConfig& cfg);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Keeps `store` resident for the server (pinned: it refuses destruction
    // until evicted) and grows the pass buffers for it; returns its id.
    // Waits for the passes in flight first.
    int admit(const DeviceStore* store);
    // Fails the store's queued requests; passes already launched finish.
    void evict(int id);
    std::uint64_t submit(int id, const double* x, std::int64_t k, double* y);
    // kDone, kFailed, or kPending when the timeout (seconds; < 0: none)
    // passed first.  A finished ticket is forgotten once waited for.
    int wait(std::uint64_t ticket, double timeout_s);
    ServerStats stats();
    void reset_stats();
    // Closed-loop sessions (see qh_server_closed_loop).
    LoadResult closed_loop(const std::vector<int>& ids, const std::vector<std::int64_t>& ks,
                           double rate, double duration, int session, std::uint64_t seed);

  private:
    struct Request {
        int store;
        const double* x;
        double* y;
        std::int64_t k;
        std::uint64_t ticket;
        double t_submit;
        int session;  // closed-loop session, or -1
    };
    struct Session {
        int store;
        std::int64_t k;
        const double* x;
        double* y;
        int left;  // requests still to submit
    };
    struct Resident {
        const DeviceStore* store = nullptr;
        std::uint64_t bytes = 0;
    };
    struct Slot;

    void dispatch_loop();
    void complete_loop();
    void launch_pass(Slot& slot, const DeviceStore& store, const std::vector<Request>& take,
                     std::int64_t K);
    void finish_locked(const Request& r, bool ok);
    void drain(std::unique_lock<std::mutex>& lock);

    ServerConfig cfg_;
    std::mutex mu_;
    std::condition_variable cv_;       // queue, slots, passes in flight
    std::condition_variable done_cv_;  // tickets and sessions
    std::deque<Request> queue_;
    std::vector<Resident> stores_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::unordered_map<std::uint64_t, int> tickets_;
    std::vector<Session> sessions_;
    int open_sessions_ = 0;
    std::uint64_t load_errors_ = 0;
    std::uint64_t next_ticket_ = 1;
    std::uint64_t resident_bytes_ = 0, buffer_bytes_ = 0;
    int inflight_ = 0;
    bool running_ = false, paused_ = false, dispatching_ = false;
    std::vector<float> latencies_ms_;
    std::uint64_t completed_ = 0, completed_rhs_ = 0, passes_ = 0;
    std::thread dispatcher_, completer_;
};

}  // namespace qhamd
