// er_server.h -- continuous-batching server for grid and Henry requests.
//
// Requests arrive at any time (er_server_submit).  A scheduler thread admits
// them in arrival order while the HBM budget allows (output buffers, Widom
// binning, and the framework unless it is already resident), then keeps up
// to `streams` launches of the tiled energy kernel in flight; every launch
// takes pending work items (grid tiles, Widom bins) from all admitted
// requests, oldest first, up to a work budget, so requests that arrive while
// others run join the next launch and their work overlaps on the device.
//
// Frameworks stay resident: uploads are kept in a content-keyed cache (atoms,
// cell, replication, cutoff), shared by every request that names the same
// framework and reference-counted; a framework no request uses stays in HBM
// until admission needs its memory (least recently used first) or the idle
// cache exceeds its byte limit.
//
// A launch computes one pair-math variant: FP64 grids (MATH_FP64) or
// everything else (MATH_MIXED).  Each launch takes the variant of the oldest
// ready request's pending work, so every request keeps making progress under
// continuous arrivals.
//
// A request completes when all its items have run (plus the Widom
// reduction); its buffers are then released (the grid stays until the
// client releases it).  Results are bit-identical to the plain path: an
// item's result depends only on its own tile and framework.
#pragma once

#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "er_runtime.h"

namespace er {

struct ServerConfig {
    int streams = 3;                     // launches in flight
    long long items_per_launch = 32768;  // grid tiles + Widom bins per launch (0 = all pending)
    int max_queue = 4096;                // waiting requests beyond which submit rejects
    long long cache_bytes = -1;          // idle frameworks kept resident: -1 up to the
                                         // HBM budget, 0 none, else at most this many bytes
    int fail_launch = 0;                 // the n-th launch attempt is refused on the host
                                         // before anything is enqueued (failure-path tests)
};

enum RequestState : int { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3,
                          RQ_REJECTED = 4 };

struct RequestSpec {
    std::vector<Particle> atoms;
    double t[9];
    int rep[3];
    double rc2;
    int grid_n[3];      // all zero: no grid
    int precision;      // grid precision (0 FP32, 1 FP64)
    uint64_t n_ins;     // 0: no Widom
    double temperature;
    uint64_t seed;
};

struct RequestInfo {
    int state = RQ_QUEUED;
    int status = ST_OK;
    std::string error;
    double t_submit = 0, t_admit = 0, t_done = 0;  // seconds, steady clock
    WidomResult widom;
};

struct CacheStats {
    long long hits = 0, misses = 0, evictions = 0;
    long long resident_bytes = 0, resident_frameworks = 0;
};

class Server {
  public:
    Server(Context *ctx, const ServerConfig &cfg);
    ~Server();
    // ST_OK, or ST_NOMEM if the launch slots could not get their buffers.
    int status() const { return status_; }
    // Returns ST_OK and a ticket, or ST_CAPACITY if the queue is full, or
    // ST_NOMEM if the request alone exceeds the HBM budget (both count as
    // rejections).
    int submit(const RequestSpec &spec, uint64_t *ticket);
    // Snapshot of a request; wait_ms < 0 blocks until it is done or failed.
    int poll(uint64_t ticket, RequestInfo *out, double wait_ms);
    // Copies a finished request's grid to host memory.
    int fetch_grid(uint64_t ticket, void *host);
    // Forget a finished request (frees its grid); waits for fetches of it
    // that are in progress on other threads.
    int release(uint64_t ticket);
    void stats(long long *submitted, long long *rejected, long long *completed,
               long long *launches);
    CacheStats cache_stats();
    // HBM bytes admission may fill (the context's cap, or the device memory
    // free at creation plus the pool's reserve, less a margin).
    size_t budget() const { return cap_; }

  private:
    struct Req;
    struct Launch;
    struct FwEntry;
    void run();
    bool admit_one(Req *r, FwEntry *hit);  // hit: the resident framework, or nullptr
    void finish(Req *r);
    void fail(Req *r);  // state RQ_FAILED (status / error already set); frees its buffers
    void fail_parts(Launch *L, const std::string &why);
    FwEntry *cache_find(const Req *r);
    FwEntry *fw_acquire(Req *r, FwEntry *hit, int *st);
    void fw_release(FwEntry *e);
    bool evict_one(const FwEntry *keep);
    void trim_idle();

    Context *ctx_;
    ServerConfig cfg_;
    size_t cap_ = 0;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<Req *> queue_;
    std::vector<Req *> active_;
    std::map<uint64_t, std::unique_ptr<Req>> all_;
    std::vector<std::unique_ptr<Launch>> launches_;
    std::unordered_multimap<uint64_t, std::unique_ptr<FwEntry>> cache_;
    size_t idle_bytes_ = 0, resident_bytes_ = 0;
    uint64_t clock_ = 0;
    CacheStats cstats_;
    uint64_t next_ticket_ = 1;
    std::atomic<bool> stop_{false};
    long long n_submitted_ = 0, n_rejected_ = 0, n_completed_ = 0, n_launches_ = 0;
    long long n_attempts_ = 0;
    std::thread worker_;
    int status_ = ST_OK;
};

}  // namespace er
