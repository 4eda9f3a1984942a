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

}  // namespace

struct Server::Req {
    uint64_t ticket = 0;
    RequestSpec spec;
    RequestInfo info;
    size_t reserve = 0;           // bytes counted against the cap
    FrameworkDev *fw = nullptr;
    GridDev *grid = nullptr;
    WidomBuffers wb;
    std::vector<void *> owned;    // Widom buffers
    JobDev grid_job{}, widom_job{};
    long long grid_next = 0, widom_next = 0;  // next local item to hand out
    long long items_total = 0, items_done = 0;
    hipEvent_t ready = nullptr;   // setup (binning) done
    hipEvent_t reduced = nullptr; // Widom sums copied to host
    double *h_sums = nullptr;     // pinned, 2 doubles
    bool reducing = false;
};

struct Server::Launch {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    JobDev *h_jobs = nullptr;     // pinned
    JobDev *d_jobs = nullptr;
    unsigned long long *d_counter = nullptr;
    double4 *d_scratch = nullptr;  // region sub-lists of this launch slot
    int scratch_wgs = 0;
    std::vector<std::pair<Req *, long long>> parts;
    bool busy = false;
};

Server::Server(Context *ctx, const ServerConfig &cfg) : ctx_(ctx), cfg_(cfg) {
    if (cfg_.streams < 1) cfg_.streams = 1;
    if (cfg_.streams > 8) cfg_.streams = 8;
    (void)hipSetDevice(ctx_->device);
    for (int i = 0; i < cfg_.streams; i++) {
        std::unique_ptr<Launch> L(new Launch());
        (void)hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
        (void)hipEventCreateWithFlags(&L->done, hipEventDisableTiming);
        (void)hipHostMalloc((void **)&L->h_jobs, sizeof(JobDev) * kMaxLaunchJobs);
        L->d_jobs = (JobDev *)ctx_->pool.alloc(sizeof(JobDev) * kMaxLaunchJobs);
        L->d_counter = (unsigned long long *)ctx_->pool.alloc(sizeof(unsigned long long));
        if (!L->stream || !L->done || !L->h_jobs || !L->d_jobs || !L->d_counter ||
            scratch_ensure(ctx_, L->stream, &L->d_scratch, &L->scratch_wgs) != ST_OK)
            status_ = ST_NOMEM;
        launches_.push_back(std::move(L));
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
        delete r->fw;
        for (void *p : r->owned) ctx_->pool.free(p);
        if (r->ready) (void)hipEventDestroy(r->ready);
        if (r->reduced) (void)hipEventDestroy(r->reduced);
        if (r->h_sums) (void)hipHostFree(r->h_sums);
    }
}

size_t Server::estimate_bytes(const RequestSpec &s) const {
    // near list (same bound as framework_upload) + atoms + grid + Widom buffers
    size_t near = framework_near_capacity((int)s.atoms.size(), s.t, s.rep, s.rc2) * 32;
    size_t b = near + s.atoms.size() * 64 + (1 << 20);
    const bool grid = s.grid_n[0] || s.grid_n[1] || s.grid_n[2];
    if (grid)
        b += (size_t)s.grid_n[0] * s.grid_n[1] * s.grid_n[2] * (s.precision == 1 ? 8 : 4);
    if (s.n_ins) b += (size_t)s.n_ins * 12 + (size_t)s.n_ins / 8 + (1 << 16);
    return b;
}

int Server::submit(const RequestSpec &spec, uint64_t *ticket) {
    std::unique_ptr<Req> r(new Req());
    r->spec = spec;
    r->info.t_submit = now_s();
    r->reserve = estimate_bytes(spec);
    std::lock_guard<std::mutex> lk(mu_);
    r->ticket = next_ticket_++;
    *ticket = r->ticket;
    n_submitted_++;
    int st = ST_OK;
    if (ctx_->hbm_cap && r->reserve > ctx_->hbm_cap) {
        st = ST_NOMEM;
        r->info.error = "request exceeds the HBM cap";
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

bool Server::admit_one(Req *r) {
    // framework upload, outputs and Widom binning, on the context stream
    const RequestSpec &s = r->spec;
    int st = framework_upload(ctx_, s.atoms.data(), (int)s.atoms.size(), s.t, s.rep, s.rc2,
                              &r->fw);
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
        r->info.state = RQ_FAILED;
        r->info.status = st;
        r->info.error = last_error();
        r->info.t_done = now_s();
        delete r->grid;
        r->grid = nullptr;
        delete r->fw;
        r->fw = nullptr;
        for (void *p : r->owned) ctx_->pool.free(p);
        r->owned.clear();
        done_cv_.notify_all();
        return false;
    }
    r->info.state = RQ_RUNNING;
    return true;
}

void Server::finish(Req *r) {
    // the framework and the Widom buffers go back to the pool; the grid stays
    // until the client fetches or releases it
    delete r->fw;
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
    delete r->fw;
    r->fw = nullptr;
    delete r->grid;
    r->grid = nullptr;
    for (void *p : r->owned) ctx_->pool.free(p);
    r->owned.clear();
    r->info.state = RQ_FAILED;
    r->info.t_done = now_s();
    done_cv_.notify_all();
}

void Server::run() {
    (void)hipSetDevice(ctx_->device);
    while (!stop_) {
        bool progress = false;
        std::unique_lock<std::mutex> lk(mu_);
        // 1. retire finished launches
        for (auto &L : launches_) {
            if (!L->busy || hipEventQuery(L->done) != hipSuccess) continue;
            for (auto &pr : L->parts) pr.first->items_done += pr.second;
            L->parts.clear();
            L->busy = false;
            progress = true;
        }
        // 2. requests whose items are all done: Widom reduction, then done
        for (size_t i = 0; i < active_.size();) {
            Req *r = active_[i];
            bool drop = false;
            if (r->items_done == r->items_total) {
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
        // 3. admit waiting requests in order while the cap allows
        while (!queue_.empty()) {
            Req *r = queue_.front();
            if (ctx_->hbm_cap && ctx_->pool.in_use() + r->reserve > ctx_->hbm_cap) break;
            queue_.pop_front();
            if (admit_one(r)) active_.push_back(r);
            progress = true;
        }
        // 4. fill free launch slots with pending items, sharing the budget
        for (auto &L : launches_) {
            if (L->busy) continue;
            std::vector<Req *> ready;
            for (Req *r : active_)
                if ((r->grid_next < r->grid_job.n_items || r->widom_next < r->widom_job.n_items) &&
                    hipEventQuery(r->ready) == hipSuccess)
                    ready.push_back(r);
            if (ready.empty()) break;
            long long budget = cfg_.items_per_launch > 0 ? cfg_.items_per_launch : (1LL << 62);
            const long long share = std::max<long long>(64, budget / (long long)ready.size());
            int nj = 0;
            long long total = 0;
            for (Req *r : ready) {
                for (int kind = 0; kind < 2 && nj < kMaxLaunchJobs && budget > 0; kind++) {
                    JobDev &src = kind == 0 ? r->grid_job : r->widom_job;
                    long long &next = kind == 0 ? r->grid_next : r->widom_next;
                    const long long left = src.n_items - next;
                    if (left <= 0) continue;
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
            // One launch per math variant: FP64-grid jobs go in a launch of their
            // own (MATH_FP64); everything else runs MATH_MIXED.
            bool any64 = false, anymix = false;
            for (int q = 0; q < nj; q++) {
                if (L->h_jobs[q].kind == JOB_GRID && L->h_jobs[q].mode == OUT_F64) any64 = true;
                else anymix = true;
            }
            int math = MATH_MIXED;
            if (any64 && anymix) {
                // keep only the mixed jobs in this launch; return the others' items
                int w = 0;
                long long begin = 0;
                std::vector<std::pair<Req *, long long>> keep;
                for (int q = 0; q < nj; q++) {
                    JobDev j = L->h_jobs[q];
                    Req *owner = L->parts[q].first;
                    if (j.kind == JOB_GRID && j.mode == OUT_F64) {
                        owner->grid_next -= j.n_items;
                        continue;
                    }
                    j.item_begin = begin;
                    begin += j.n_items;
                    L->h_jobs[w++] = j;
                    keep.push_back(L->parts[q]);
                }
                nj = w;
                total = begin;
                L->parts.swap(keep);
            } else if (any64) {
                math = MATH_FP64;
            }
            (void)hipMemcpyAsync(L->d_jobs, L->h_jobs, sizeof(JobDev) * nj, hipMemcpyHostToDevice,
                                 L->stream);
            (void)launch_fill_u64(L->d_counter, 0ull, 1, L->stream);
            const int wgs = tiles_workgroups(ctx_, math, total, L->scratch_wgs);
            (void)launch_tiles(math, ctx_->tuning.ppl, L->d_jobs, nj, total, L->d_counter, wgs,
                               ctx_->tuning.cap, L->d_scratch, kRegionRecords, L->stream);
            (void)hipEventRecord(L->done, L->stream);
            L->busy = true;
            n_launches_++;
            progress = true;
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
    GridDev *g = it->second->grid;
    lk.unlock();
    if (hipMemcpy(host, g->d_data, g->bytes(), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("server: grid copy failed");
        return ST_HIP;
    }
    return ST_OK;
}

int Server::release(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
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
    delete r->grid;
    if (r->ready) (void)hipEventDestroy(r->ready);
    if (r->reduced) (void)hipEventDestroy(r->reduced);
    if (r->h_sums) (void)hipHostFree(r->h_sums);
    all_.erase(it);
    return ST_OK;
}

void Server::stats(long long *submitted, long long *rejected, long long *completed,
                   long long *launches) {
    std::lock_guard<std::mutex> lk(mu_);
    if (submitted) *submitted = n_submitted_;
    if (rejected) *rejected = n_rejected_;
    if (completed) *completed = n_completed_;
    if (launches) *launches = n_launches_;
}

}  // namespace er
