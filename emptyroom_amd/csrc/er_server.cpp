// er_server.cpp -- see er_server.h.
#include "er_server.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sstream>

namespace er {

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

constexpr int kMaxLaunchJobs = 1024;

// Hash of the bytes that define an uploaded framework: 64-bit words mixed
// with a multiply-xorshift step (a hit is confirmed by a full comparison).
uint64_t hash_bytes(uint64_t h, const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    auto mix = [&h](uint64_t w) {
        h ^= w;
        h *= 0x9E3779B97F4A7C15ULL;
        h ^= h >> 29;
    };
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        mix(w);
    }
    uint64_t t = 0;
    std::memcpy(&t, b + i, n - i);
    mix(t ^ ((uint64_t)n << 56));
    return h;
}

uint64_t framework_key(const RequestSpec &s) {
    uint64_t h = 0xcbf29ce484222325ULL;
    h = hash_bytes(h, s.atoms.data(), s.atoms.size() * sizeof(Particle));
    h = hash_bytes(h, s.t, sizeof(s.t));
    h = hash_bytes(h, s.rep, sizeof(s.rep));
    h = hash_bytes(h, &s.rc2, sizeof(s.rc2));
    return h;
}

// Device bytes of a framework upload (near list + atoms + tables), an upper bound.
size_t framework_estimate(const RequestSpec &s) {
    const size_t near = framework_near_capacity((int)s.atoms.size(), s.t, s.rep, s.rc2) * 32;
    return near + s.atoms.size() * 64 + (1 << 20);
}

// Device bytes of a request's outputs and Widom buffers.
size_t work_estimate(const RequestSpec &s) {
    size_t b = 1 << 16;
    const bool grid = s.grid_n[0] || s.grid_n[1] || s.grid_n[2];
    if (grid)
        b += (size_t)s.grid_n[0] * s.grid_n[1] * s.grid_n[2] * (s.precision == 1 ? 8 : 4);
    if (s.n_ins) b += (size_t)s.n_ins * 12 + (size_t)s.n_ins / 8 + (1 << 16);
    return b;
}

bool is_fp64_grid(const JobDev &j) { return j.kind == JOB_GRID && j.mode == OUT_F64; }

}  // namespace

struct Server::FwEntry {
    uint64_t key = 0;
    std::vector<Particle> atoms;
    double t[9];
    int rep[3];
    double rc2 = 0;
    FrameworkDev *fw = nullptr;
    size_t bytes = 0;       // pool bytes the upload took
    int refs = 0;
    uint64_t last_use = 0;
};

struct Server::Req {
    uint64_t ticket = 0;
    uint64_t key = 0;             // framework content key
    RequestSpec spec;
    RequestInfo info;
    size_t fw_reserve = 0;        // bytes counted against the budget for the framework
    size_t work_reserve = 0;      // ... for outputs and Widom buffers
    FwEntry *fwe = nullptr;
    FrameworkDev *fw = nullptr;
    GridDev *grid = nullptr;
    WidomBuffers wb;
    std::vector<void *> owned;    // Widom buffers
    JobDev grid_job{}, widom_job{};
    long long grid_next = 0, widom_next = 0;  // next local item to hand out
    long long items_total = 0, items_done = 0;
    int inflight = 0;             // launch parts of this request not yet retired
    bool failed = false;          // fails once nothing of it is in flight
    int fetching = 0;             // fetch_grid copies in progress
    hipEvent_t ready = nullptr;   // setup (binning) done
    hipEvent_t reduced = nullptr; // Widom sums copied to host
    double *h_sums = nullptr;     // pinned, 2 doubles
    bool reducing = false;
    bool is_ready = false;        // `ready` seen complete (queried until then only)
};

struct Server::Launch {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    JobDev *h_jobs = nullptr;     // pinned
    JobDev *d_jobs = nullptr;
    unsigned long long *d_counter = nullptr;
    uint32_t *d_scratch = nullptr;  // region sub-lists of this launch slot
    int scratch_wgs = 0;
    std::vector<std::pair<Req *, long long>> parts;
    bool busy = false;
};

Server::Server(Context *ctx, const ServerConfig &cfg) : ctx_(ctx), cfg_(cfg) {
    if (cfg_.streams < 1) cfg_.streams = 1;
    if (cfg_.streams > 8) cfg_.streams = 8;
    if (hipSetDevice(ctx_->device) != hipSuccess) {
        status_ = ST_HIP;
        return;
    }
    for (int i = 0; i < cfg_.streams; i++) {
        std::unique_ptr<Launch> L(new Launch());
        if (hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess)
            L->stream = nullptr;
        if (hipEventCreateWithFlags(&L->done, hipEventDisableTiming) != hipSuccess)
            L->done = nullptr;
        if (hipHostMalloc((void **)&L->h_jobs, sizeof(JobDev) * kMaxLaunchJobs) != hipSuccess)
            L->h_jobs = nullptr;
        L->d_jobs = (JobDev *)ctx_->pool.alloc(sizeof(JobDev) * kMaxLaunchJobs);
        L->d_counter = (unsigned long long *)ctx_->pool.alloc(sizeof(unsigned long long));
        if (!L->stream || !L->done || !L->h_jobs || !L->d_jobs || !L->d_counter ||
            scratch_ensure(ctx_, L->stream, &L->d_scratch, &L->scratch_wgs) != ST_OK)
            status_ = ST_NOMEM;
        launches_.push_back(std::move(L));
    }
    (void)hipGetLastError();
    // The admission budget: the context's cap, or else what the device holds
    // now (free memory plus the pool's reserve) less a margin, so that an
    // over-subscribed request waits instead of failing at hipMalloc.
    if (ctx_->hbm_cap) {
        cap_ = ctx_->hbm_cap;
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            status_ = ST_HIP;
        }
        const size_t margin = std::max<size_t>((size_t)1 << 30, tot / 20);
        cap_ = ctx_->pool.reserved() + (fr > margin ? fr - margin : 0);
    }
    if (status_ == ST_OK) worker_ = std::thread([this] { run(); });
}

Server::~Server() {
    stop_ = true;
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipStreamSynchronize(ctx_->stream);  // uploads and Widom binning
    for (auto &L : launches_) {
        if (L->stream) {
            (void)hipStreamSynchronize(L->stream);
            (void)hipStreamDestroy(L->stream);
        }
        if (L->done) (void)hipEventDestroy(L->done);
        if (L->h_jobs) (void)hipHostFree(L->h_jobs);
        ctx_->pool.free(L->d_jobs);
        ctx_->pool.free(L->d_counter);
        ctx_->pool.free(L->d_scratch);
    }
    for (auto &kv : all_) {
        Req *r = kv.second.get();
        delete r->grid;
        for (void *p : r->owned) ctx_->pool.free(p);
        if (r->ready) (void)hipEventDestroy(r->ready);
        if (r->reduced) (void)hipEventDestroy(r->reduced);
        if (r->h_sums) (void)hipHostFree(r->h_sums);
    }
    for (auto &kv : cache_) delete kv.second->fw;
}

// ---------------------------------------------------------------------------
// Framework cache (called with mu_ held, on the scheduler thread)
// ---------------------------------------------------------------------------

Server::FwEntry *Server::cache_find(const Req *r) {
    auto range = cache_.equal_range(r->key);
    for (auto it = range.first; it != range.second; ++it) {
        FwEntry *e = it->second.get();
        const RequestSpec &s = r->spec;
        if (e->atoms.size() == s.atoms.size() && e->rc2 == s.rc2 &&
            std::memcmp(e->t, s.t, sizeof(e->t)) == 0 &&
            std::memcmp(e->rep, s.rep, sizeof(e->rep)) == 0 &&
            std::memcmp(e->atoms.data(), s.atoms.data(), s.atoms.size() * sizeof(Particle)) == 0)
            return e;
    }
    return nullptr;
}

Server::FwEntry *Server::fw_acquire(Req *r, FwEntry *e, int *st) {
    if (e) {
        if (e->refs == 0) idle_bytes_ -= e->bytes;
        e->refs++;
        e->last_use = ++clock_;
        cstats_.hits++;
        *st = ST_OK;
        return e;
    }
    cstats_.misses++;
    const RequestSpec &s = r->spec;
    const size_t before = ctx_->pool.in_use();
    FrameworkDev *fw = nullptr;
    *st = framework_upload(ctx_, s.atoms.data(), (int)s.atoms.size(), s.t, s.rep, s.rc2, &fw);
    if (*st != ST_OK) return nullptr;
    std::unique_ptr<FwEntry> ne(new FwEntry());
    ne->key = r->key;
    ne->atoms = s.atoms;
    std::memcpy(ne->t, s.t, sizeof(ne->t));
    std::memcpy(ne->rep, s.rep, sizeof(ne->rep));
    ne->rc2 = s.rc2;
    ne->fw = fw;
    const size_t after = ctx_->pool.in_use();
    ne->bytes = after > before ? after - before : 0;
    ne->refs = 1;
    ne->last_use = ++clock_;
    resident_bytes_ += ne->bytes;
    e = ne.get();
    cache_.emplace(r->key, std::move(ne));
    return e;
}

void Server::fw_release(FwEntry *e) {
    if (!e || e->refs <= 0) return;
    if (--e->refs == 0) {
        e->last_use = ++clock_;
        idle_bytes_ += e->bytes;
        trim_idle();
    }
}

// Evicts the least recently used framework no request holds (other than
// keep); false if there is none.
bool Server::evict_one(const FwEntry *keep) {
    auto victim = cache_.end();
    for (auto it = cache_.begin(); it != cache_.end(); ++it) {
        const FwEntry *e = it->second.get();
        if (e->refs != 0 || e == keep) continue;
        if (victim == cache_.end() || e->last_use < victim->second->last_use) victim = it;
    }
    if (victim == cache_.end()) return false;
    FwEntry *e = victim->second.get();
    idle_bytes_ -= e->bytes;
    resident_bytes_ -= e->bytes;
    delete e->fw;  // no launch or upload uses an unreferenced framework
    cache_.erase(victim);
    cstats_.evictions++;
    return true;
}

void Server::trim_idle() {
    if (cfg_.cache_bytes < 0) return;
    while (idle_bytes_ > (size_t)cfg_.cache_bytes && evict_one(nullptr)) {
    }
}

// ---------------------------------------------------------------------------
// Requests
// ---------------------------------------------------------------------------

int Server::submit(const RequestSpec &spec, uint64_t *ticket) {
    std::unique_ptr<Req> r(new Req());
    r->spec = spec;
    r->info.t_submit = now_s();
    r->key = framework_key(spec);
    r->fw_reserve = framework_estimate(spec);
    r->work_reserve = work_estimate(spec);
    std::lock_guard<std::mutex> lk(mu_);
    r->ticket = next_ticket_++;
    *ticket = r->ticket;
    n_submitted_++;
    int st = ST_OK;
    if (r->fw_reserve + r->work_reserve > cap_) {
        st = ST_NOMEM;
        r->info.error = "request exceeds the HBM budget";
    } else if ((int)queue_.size() >= cfg_.max_queue) {
        st = ST_CAPACITY;
        r->info.error = "server queue full";
    }
    if (st != ST_OK) {
        r->info.state = RQ_REJECTED;
        r->info.status = st;
        r->info.t_done = r->info.t_submit;
        n_rejected_++;
        all_[r->ticket] = std::move(r);
        done_cv_.notify_all();
        return st;
    }
    queue_.push_back(r.get());
    all_[r->ticket] = std::move(r);
    cv_.notify_all();
    return ST_OK;
}

bool Server::admit_one(Req *r, FwEntry *hit) {
    // framework (resident or uploaded), outputs and Widom binning, on the
    // context stream
    const RequestSpec &s = r->spec;
    int st = ST_OK;
    r->fwe = fw_acquire(r, hit, &st);
    if (r->fwe) r->fw = r->fwe->fw;
    const bool grid = s.grid_n[0] || s.grid_n[1] || s.grid_n[2];
    if (st == ST_OK && grid) {
        r->grid = new GridDev();
        r->grid->ctx = ctx_;
        r->grid->fw = r->fw;
        std::memcpy(r->grid->n, s.grid_n, sizeof(r->grid->n));
        r->grid->precision = s.precision;
        r->grid->d_data = ctx_->pool.alloc(r->grid->bytes());
        if (!r->grid->d_data) st = ST_NOMEM;
    }
    if (st == ST_OK && grid) {
        const int math = s.precision == 1 ? MATH_FP64 : MATH_MIXED;
        r->grid_job = make_grid_job(r->fw, s.grid_n, s.precision, r->grid->d_data, ctx_->tuning,
                                    math);
        r->items_total += r->grid_job.n_items;
    }
    if (st == ST_OK && s.n_ins) {
        st = widom_prepare(ctx_, r->fw, s.n_ins, &r->wb, &r->owned);
        if (st == ST_OK) {
            r->widom_job = make_widom_job(r->fw, s.n_ins, s.seed, r->wb, ctx_->tuning,
                                          kMinGridItems, s.temperature);
            r->items_total += r->widom_job.n_items;
            st = widom_enqueue_binning(ctx_, r->wb, s.n_ins, s.seed, 0u, ctx_->stream);
        }
        if (st == ST_OK && hipHostMalloc((void **)&r->h_sums, 2 * sizeof(double)) != hipSuccess)
            st = ST_NOMEM;
        if (st == ST_OK && hipEventCreateWithFlags(&r->reduced, hipEventDisableTiming) !=
                               hipSuccess)
            st = ST_HIP;
    }
    if (st == ST_OK && (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) != hipSuccess ||
                        hipEventRecord(r->ready, ctx_->stream) != hipSuccess))
        st = ST_HIP;
    r->info.t_admit = now_s();
    if (st != ST_OK) {
        (void)hipGetLastError();
        // binning may already be enqueued on the context stream
        (void)hipStreamSynchronize(ctx_->stream);
        r->info.status = st;
        r->info.error = last_error();
        fail(r);
        return false;
    }
    r->info.state = RQ_RUNNING;
    return true;
}

void Server::finish(Req *r) {
    // the Widom buffers go back to the pool and the framework back to the
    // cache; the grid stays until the client releases it
    fw_release(r->fwe);
    r->fwe = nullptr;
    r->fw = nullptr;
    if (r->grid) r->grid->fw = nullptr;
    for (void *p : r->owned) ctx_->pool.free(p);
    r->owned.clear();
    if (r->h_sums) {
        r->info.widom.sum_b = r->h_sums[0];
        r->info.widom.sum_eb = r->h_sums[1];
        r->info.widom.n = r->spec.n_ins;
    }
    r->info.state = RQ_DONE;
    r->info.t_done = now_s();
    n_completed_++;
    done_cv_.notify_all();
}

void Server::fail(Req *r) {
    fw_release(r->fwe);
    r->fwe = nullptr;
    r->fw = nullptr;
    delete r->grid;
    r->grid = nullptr;
    for (void *p : r->owned) ctx_->pool.free(p);
    r->owned.clear();
    r->info.state = RQ_FAILED;
    r->info.t_done = now_s();
    done_cv_.notify_all();
}

// A launch that could not be enqueued in full: whatever part of it was
// enqueued has finished once the stream is synchronized, and the requests
// whose items it held fail (their other launches are left to retire).
void Server::fail_parts(Launch *L, const std::string &why) {
    (void)hipStreamSynchronize(L->stream);
    (void)hipGetLastError();
    for (auto &pr : L->parts) {
        Req *o = pr.first;
        if (!o->failed) {
            o->failed = true;
            o->info.status = ST_HIP;
            o->info.error = why;
        }
    }
    L->parts.clear();
    L->busy = false;
}

void Server::run() {
    if (hipSetDevice(ctx_->device) != hipSuccess) (void)hipGetLastError();
    while (!stop_) {
        bool progress = false;
        std::unique_lock<std::mutex> lk(mu_);
        // 1. retire finished launches
        for (auto &L : launches_) {
            if (!L->busy || hipEventQuery(L->done) != hipSuccess) continue;
            for (auto &pr : L->parts) {
                pr.first->items_done += pr.second;
                pr.first->inflight--;
            }
            L->parts.clear();
            L->busy = false;
            progress = true;
        }
        // 2. failed requests with nothing in flight fail; requests whose items
        //    are all done get their Widom reduction, then are done
        for (size_t i = 0; i < active_.size();) {
            Req *r = active_[i];
            bool drop = false;
            if (r->failed) {
                if (r->inflight == 0) {
                    fail(r);
                    drop = true;
                    progress = true;
                }
            } else if (r->items_done == r->items_total) {
                if (r->spec.n_ins && !r->reducing) {
                    hipStream_t s = launches_[0]->stream;
                    double2 *part = nullptr;
                    double *sums = nullptr;
                    part = (double2 *)ctx_->pool.alloc(sizeof(double2) * kReduceBlocks);
                    sums = (double *)ctx_->pool.alloc(2 * sizeof(double));
                    if (part) r->owned.push_back(part);
                    if (sums) r->owned.push_back(sums);
                    hipError_t e = hipErrorOutOfMemory;
                    if (part && sums) {
                        e = launch_widom_partials(r->wb.e, (long long)r->spec.n_ins,
                                                  r->spec.temperature, part, s);
                        if (e == hipSuccess) e = launch_widom_final(part, sums, s);
                        if (e == hipSuccess)
                            e = hipMemcpyAsync(r->h_sums, sums, 2 * sizeof(double),
                                               hipMemcpyDeviceToHost, s);
                        if (e == hipSuccess) e = hipEventRecord(r->reduced, s);
                    }
                    if (e != hipSuccess) {
                        // no complete reduction in flight: the request fails
                        // (an event that was never recorded would read as
                        // complete); wait for any part already enqueued
                        (void)hipStreamSynchronize(s);
                        (void)hipGetLastError();
                        r->info.status = e == hipErrorOutOfMemory ? ST_NOMEM : ST_HIP;
                        r->info.error = "Widom reduction could not be enqueued";
                        fail(r);
                        drop = true;
                    }
                    r->reducing = true;
                    progress = true;
                } else if (!r->spec.n_ins || hipEventQuery(r->reduced) == hipSuccess) {
                    finish(r);
                    drop = true;
                    progress = true;
                }
            }
            if (drop) active_.erase(active_.begin() + i);
            else i++;
        }
        // 3. admit waiting requests in order while the budget allows; a
        //    resident framework costs nothing, idle ones are evicted for room
        while (!queue_.empty()) {
            Req *r = queue_.front();
            FwEntry *hit = cache_find(r);
            const size_t need = r->work_reserve + (hit ? 0 : r->fw_reserve);
            bool fits = ctx_->pool.in_use() + need <= cap_;
            while (!fits && evict_one(hit)) fits = ctx_->pool.in_use() + need <= cap_;
            if (!fits) break;
            queue_.pop_front();
            if (admit_one(r, hit)) active_.push_back(r);
            progress = true;
        }
        // 4. fill free launch slots with pending items
        for (auto &L : launches_) {
            if (L->busy) continue;
            std::vector<Req *> ready;
            for (Req *r : active_)
                if (!r->failed &&
                    (r->grid_next < r->grid_job.n_items || r->widom_next < r->widom_job.n_items) &&
                    (r->is_ready || (r->is_ready = hipEventQuery(r->ready) == hipSuccess)))
                    ready.push_back(r);
            if (ready.empty()) break;
            // the pair-math variant of this launch: that of the oldest ready
            // request's pending work (FP64 grid first)
            const Req *o = ready.front();
            const bool fp64 = is_fp64_grid(o->grid_job) && o->grid_next < o->grid_job.n_items;
            const int math = fp64 ? MATH_FP64 : MATH_MIXED;
            auto job_math = [](const JobDev &j) { return is_fp64_grid(j) ? MATH_FP64 : MATH_MIXED; };
            long long nkind = 0;
            for (Req *r : ready)
                for (int kind = 0; kind < 2; kind++) {
                    const JobDev &src = kind == 0 ? r->grid_job : r->widom_job;
                    const long long next = kind == 0 ? r->grid_next : r->widom_next;
                    if (src.n_items > next && job_math(src) == math) nkind++;
                }
            long long budget = cfg_.items_per_launch > 0 ? cfg_.items_per_launch : (1LL << 62);
            const long long share = std::max<long long>(64, budget / std::max(1LL, nkind));
            int nj = 0;
            long long total = 0;
            for (Req *r : ready) {
                for (int kind = 0; kind < 2 && nj < kMaxLaunchJobs && budget > 0; kind++) {
                    JobDev &src = kind == 0 ? r->grid_job : r->widom_job;
                    long long &next = kind == 0 ? r->grid_next : r->widom_next;
                    const long long left = src.n_items - next;
                    if (left <= 0 || job_math(src) != math) continue;
                    // a grid item is a region of tiles, a Widom item a region of bins
                    const long long w = (long long)src.region[0] * src.region[1] * src.region[2];
                    const long long take =
                        std::min(left, std::max<long long>(1, std::min(share, budget) / w));
                    JobDev j = src;
                    j.local0 = next;
                    j.n_items = take;
                    j.item_begin = total;
                    L->h_jobs[nj++] = j;
                    L->parts.push_back({r, take});
                    next += take;
                    total += take;
                    budget -= take * w;
                }
            }
            if (nj == 0) break;
            progress = true;
            n_attempts_++;
            hipError_t e = hipSuccess;
            std::string why;
            if (cfg_.fail_launch > 0 && n_attempts_ == cfg_.fail_launch) {
                e = hipErrorInvalidValue;
                why = "launch refused on the host (fail_launch)";
            }
            if (e == hipSuccess) {
                e = hipMemcpyAsync(L->d_jobs, L->h_jobs, sizeof(JobDev) * nj,
                                   hipMemcpyHostToDevice, L->stream);
                if (e != hipSuccess) why = "job table copy failed";
            }
            if (e == hipSuccess) {
                e = launch_fill_u64(L->d_counter, 0ull, 1, L->stream);
                if (e != hipSuccess) why = "counter fill failed";
            }
            if (e == hipSuccess) {
                const int wgs = tiles_workgroups(ctx_, math, total, L->scratch_wgs);
                e = launch_tiles(math, ctx_->tuning.ppl, L->d_jobs, nj, total, L->d_counter, wgs,
                                 ctx_->tuning.cap, L->d_scratch, kRegionRecords, L->stream);
                if (e != hipSuccess) why = std::string("tile kernel launch failed: ") +
                                           hipGetErrorString(e);
            }
            if (e == hipSuccess) {
                e = hipEventRecord(L->done, L->stream);
                if (e != hipSuccess) why = "launch event record failed";
            }
            if (e != hipSuccess) {
                fail_parts(L.get(), why);
                continue;
            }
            for (auto &pr : L->parts) pr.first->inflight++;
            L->busy = true;
            n_launches_++;
        }
        if (!progress) {
            bool inflight = false;
            for (auto &L : launches_) inflight |= L->busy;
            if (inflight || !active_.empty())
                cv_.wait_for(lk, std::chrono::microseconds(50));
            else
                cv_.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

int Server::poll(uint64_t ticket, RequestInfo *out, double wait_ms) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(ticket);
    if (it == all_.end()) {
        set_error("server: unknown ticket");
        return ST_ARG;
    }
    Req *r = it->second.get();
    auto finished = [r] { return r->info.state >= RQ_DONE; };
    if (wait_ms < 0) done_cv_.wait(lk, finished);
    else if (wait_ms > 0)
        done_cv_.wait_for(lk, std::chrono::microseconds((long long)(wait_ms * 1000)), finished);
    *out = r->info;
    return ST_OK;
}

int Server::fetch_grid(uint64_t ticket, void *host) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(ticket);
    if (it == all_.end() || it->second->info.state != RQ_DONE || !it->second->grid) {
        set_error("server: no finished grid for this ticket");
        return ST_STATE;
    }
    Req *r = it->second.get();
    GridDev *g = r->grid;
    r->fetching++;  // release() waits until this copy is over
    lk.unlock();
    const hipError_t e = hipMemcpy(host, g->d_data, g->bytes(), hipMemcpyDeviceToHost);
    lk.lock();
    r->fetching--;
    done_cv_.notify_all();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("server: grid copy failed");
        return ST_HIP;
    }
    return ST_OK;
}

int Server::release(uint64_t ticket) {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        auto it = all_.find(ticket);
        if (it == all_.end()) {
            set_error("server: unknown ticket");
            return ST_ARG;
        }
        Req *r = it->second.get();
        if (r->info.state < RQ_DONE) {
            set_error("server: request still running");
            return ST_STATE;
        }
        if (r->fetching > 0) {
            done_cv_.wait(lk);
            continue;  // look the ticket up again: another thread may have released it
        }
        delete r->grid;
        if (r->ready) (void)hipEventDestroy(r->ready);
        if (r->reduced) (void)hipEventDestroy(r->reduced);
        if (r->h_sums) (void)hipHostFree(r->h_sums);
        all_.erase(it);
        return ST_OK;
    }
}

void Server::stats(long long *submitted, long long *rejected, long long *completed,
                   long long *launches) {
    std::lock_guard<std::mutex> lk(mu_);
    if (submitted) *submitted = n_submitted_;
    if (rejected) *rejected = n_rejected_;
    if (completed) *completed = n_completed_;
    if (launches) *launches = n_launches_;
}

CacheStats Server::cache_stats() {
    std::lock_guard<std::mutex> lk(mu_);
    CacheStats s = cstats_;
    s.resident_bytes = (long long)resident_bytes_;
    s.resident_frameworks = (long long)cache_.size();
    return s;
}

}  // namespace er
