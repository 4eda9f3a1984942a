// er_capi.cpp -- extern "C" entry points declared in include/emptyroom.h.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/emptyroom.h"
#include "er_host.h"
#include "er_runtime.h"
#include "er_server.h"

using er::ST_ARG;
using er::ST_IO;
using er::ST_OK;

struct er_host_run {
    int mode = 0;
    er::host::GridParams gp;
    er::host::HenryParams hp;
    er::host::RunSetup setup;
};

namespace {

er::Context *C(er_context *c) { return reinterpret_cast<er::Context *>(c); }
const er::Context *C(const er_context *c) { return reinterpret_cast<const er::Context *>(c); }
er::FrameworkDev *F(er_framework *f) { return reinterpret_cast<er::FrameworkDev *>(f); }
er::GridDev *G(er_grid *g) { return reinterpret_cast<er::GridDev *>(g); }
const er::GridDev *G(const er_grid *g) { return reinterpret_cast<const er::GridDev *>(g); }
er::Batch *B(er_batch *b) { return reinterpret_cast<er::Batch *>(b); }

int arg_error(const char *msg) {
    er::set_error(msg);
    return ST_ARG;
}

template <class Fn>
int guarded(Fn fn) {
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        er::set_error("host allocation failed");
        return er::ST_NOMEM;
    } catch (...) {
        er::set_error("unexpected exception inside libemptyroom");
        return er::ST_STATE;
    }
}

}  // namespace

extern "C" {

const char *er_last_error(void) { return er::last_error(); }
const char *er_version(void) { return "emptyroom-mi355x 0.1 (gfx950)"; }

int er_context_create(int device, size_t hbm_cap_bytes, er_context **out) {
    if (!out) return arg_error("er_context_create: out is null");
    return guarded([&]() -> int {
        er::Context *c = nullptr;
        int st = er::context_create(device, hbm_cap_bytes, &c);
        if (st == ST_OK) *out = reinterpret_cast<er_context *>(c);
        return st;
    });
}

int er_context_destroy(er_context *ctx) {
    delete C(ctx);
    return ST_OK;
}

}  // extern "C"

namespace {
void tuning_to(const er::Tuning &t, er_tuning *out) {
    out->points_per_lane = t.ppl;
    out->lds_records = t.cap;
    for (int d = 0; d < 3; d++) out->tile[d] = t.tile[d];
    out->widom_per_bin = t.widom_per_bin;
    out->s2_switch = t.s2_switch;
    out->region_len = t.region_len;
    for (int c = 0; c < 2; c++) out->class_spacing[c] = t.class_spacing[c];
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) out->class_tile[c][d] = t.class_tile[c][d];
}
}  // namespace

extern "C" {

int er_context_get_tuning(const er_context *ctx, er_tuning *out) {
    if (!ctx || !out) return arg_error("er_context_get_tuning: null argument");
    tuning_to(C(ctx)->tuning, out);
    return ST_OK;
}

int er_tuning_defaults(er_tuning *out) {
    if (!out) return arg_error("er_tuning_defaults: null argument");
    tuning_to(er::Tuning(), out);
    return ST_OK;
}

}  // extern "C"

namespace {
er::Tuning tuning_from(const er_tuning *in) {
    er::Tuning t;
    t.ppl = in->points_per_lane;
    t.cap = in->lds_records;
    for (int d = 0; d < 3; d++) t.tile[d] = in->tile[d];
    t.widom_per_bin = in->widom_per_bin;
    t.s2_switch = in->s2_switch;
    t.region_len = in->region_len;
    for (int c = 0; c < 2; c++) t.class_spacing[c] = in->class_spacing[c];
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) t.class_tile[c][d] = in->class_tile[c][d];
    return t;
}
}  // namespace

extern "C" {

int er_grid_tile_for(const er_tuning *tuning, const double t_matrix[9], const int n[3],
                     int tile_out[3], int *size_class) {
    if (!tuning || !t_matrix || !n || !tile_out)
        return arg_error("er_grid_tile_for: null argument");
    const er::Tuning t = tuning_from(tuning);
    if (er::check_tuning(t) != ST_OK) return ST_ARG;
    for (int d = 0; d < 3; d++)
        if (n[d] < 2) return arg_error("er_grid_tile_for: grid dimensions must be >= 2");
    er::grid_tile_shape(t_matrix, n, t, tile_out);
    if (size_class) *size_class = er::grid_size_class(t_matrix, n, t);
    return ST_OK;
}

int er_context_set_tuning(er_context *ctx, const er_tuning *in) {
    if (!ctx || !in) return arg_error("er_context_set_tuning: null argument");
    if (in->points_per_lane != 1 && in->points_per_lane != 2 && in->points_per_lane != 4)
        return arg_error("er_context_set_tuning: points_per_lane must be 1, 2 or 4");
    if (in->tile[0] < 1 || in->tile[1] < 1 || in->tile[2] < 1 ||
        in->tile[0] * in->tile[1] * in->tile[2] != 256 * in->points_per_lane)
        return arg_error("er_context_set_tuning: tile must multiply to 256 * points_per_lane");
    if (in->lds_records < 512 || in->lds_records > 4800)
        return arg_error("er_context_set_tuning: lds_records must be in [512, 4800]");
    if (in->widom_per_bin < 1) return arg_error("er_context_set_tuning: widom_per_bin < 1");
    if (!(in->s2_switch > 0.0) || !(in->s2_switch <= 1.0))
        return arg_error("er_context_set_tuning: s2_switch must be in (0, 1]");
    if (!(in->region_len >= 0.0) || !(in->region_len <= 1000.0))
        return arg_error("er_context_set_tuning: region_len must be in [0, 1000] Angstrom");
    const er::Tuning nt = tuning_from(in);
    if (er::check_tuning(nt) != ST_OK) return ST_ARG;
    er::Context *c = C(ctx);
    c->tuning = nt;
    // a tuning with higher occupancy needs more region scratch slices
    return guarded([&]() -> int {
        return er::scratch_ensure(c, c->stream, &c->d_scratch, &c->scratch_wgs);
    });
}

int er_context_memory(const er_context *ctx, size_t *in_use, size_t *reserved) {
    if (!ctx) return arg_error("er_context_memory: null context");
    er::Context *c = const_cast<er::Context *>(C(ctx));
    if (in_use) *in_use = c->pool.in_use();
    if (reserved) *reserved = c->pool.reserved();
    return ST_OK;
}

int er_framework_upload(er_context *ctx, const er_particle *atoms, int n_atoms,
                        const double t_matrix[9], const int rep[3], double rc2,
                        er_framework **out) {
    if (!ctx || !out) return arg_error("er_framework_upload: null argument");
    static_assert(sizeof(er_particle) == sizeof(er::Particle), "particle layout");
    return guarded([&]() -> int {
        er::FrameworkDev *f = nullptr;
        int st = er::framework_upload(C(ctx), reinterpret_cast<const er::Particle *>(atoms),
                                      n_atoms, t_matrix, rep, rc2, &f);
        if (st == ST_OK) *out = reinterpret_cast<er_framework *>(f);
        return st;
    });
}

int er_framework_release(er_framework *fw) {
    delete F(fw);
    return ST_OK;
}

int er_grid_compute(er_context *ctx, er_framework *fw, const int n[3], int precision,
                    void *host_out, er_grid **resident_out) {
    if (!ctx || !fw || !n) return arg_error("er_grid_compute: null argument");
    return guarded([&]() -> int {
        er::GridDev *g = nullptr;
        int st = er::grid_compute(C(ctx), F(fw), n, precision, host_out,
                                  resident_out ? &g : nullptr);
        if (st == ST_OK && resident_out) *resident_out = reinterpret_cast<er_grid *>(g);
        return st;
    });
}

int er_grid_download(er_grid *grid, void *host_out) {
    if (!grid || !host_out) return arg_error("er_grid_download: null argument");
    er::GridDev *g = G(grid);
    if (hipSetDevice(g->ctx->device) != hipSuccess ||
        hipMemcpyAsync(host_out, g->d_data, g->bytes(), hipMemcpyDeviceToHost, g->ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(g->ctx->stream) != hipSuccess) {
        er::set_error("er_grid_download: copy failed");
        return er::ST_HIP;
    }
    return ST_OK;
}

int er_grid_info(const er_grid *grid, int n[3], int *precision) {
    if (!grid) return arg_error("er_grid_info: null grid");
    const er::GridDev *g = G(grid);
    if (n)
        for (int d = 0; d < 3; d++) n[d] = g->n[d];
    if (precision) *precision = g->precision;
    return ST_OK;
}

int er_grid_release(er_grid *grid) {
    delete G(grid);
    return ST_OK;
}

int er_widom(er_context *ctx, er_framework *fw, uint64_t n_insertions, double temperature,
             uint64_t seed, er_widom_result *out, double *energies_out) {
    if (!ctx || !fw || !out) return arg_error("er_widom: null argument");
    return guarded([&]() -> int {
        er::WidomResult r;
        int st = er::widom_run(C(ctx), F(fw), n_insertions, temperature, seed, &r, energies_out);
        if (st == ST_OK) {
            out->sum_b = r.sum_b;
            out->sum_eb = r.sum_eb;
            out->n = r.n;
        }
        return st;
    });
}

int er_count_incut_grid(er_context *ctx, er_framework *fw, const int n[3], uint64_t *out) {
    if (!ctx || !fw || !n || !out) return arg_error("er_count_incut_grid: null argument");
    return guarded([&]() -> int {
        unsigned long long v = 0;
        int st = er::count_incut_grid(C(ctx), F(fw), n, &v);
        *out = v;
        return st;
    });
}

int er_grid_compute_noalloc(er_context *ctx, er_framework *fw, const int n[3], int precision,
                            void *device_out) {
    if (!ctx || !fw || !n || !device_out)
        return arg_error("er_grid_compute_noalloc: null argument");
    return guarded([&]() -> int {
        return er::grid_compute_noalloc(C(ctx), F(fw), n, precision, device_out);
    });
}

int er_count_incut_grid_map(er_context *ctx, er_framework *fw, const int n[3], uint64_t *total,
                            uint32_t *counts_out) {
    if (!ctx || !fw || !n || !counts_out)
        return arg_error("er_count_incut_grid_map: null argument");
    return guarded([&]() -> int {
        unsigned long long v = 0;
        int st = er::count_incut_grid(C(ctx), F(fw), n, &v, counts_out);
        if (total) *total = v;
        return st;
    });
}

int er_count_incut_widom(er_context *ctx, er_framework *fw, uint64_t n_insertions, uint64_t seed,
                         uint64_t *out) {
    if (!ctx || !fw || !out) return arg_error("er_count_incut_widom: null argument");
    return guarded([&]() -> int {
        unsigned long long v = 0;
        int st = er::count_incut_widom(C(ctx), F(fw), n_insertions, seed, &v);
        *out = v;
        return st;
    });
}

void er_widom_point(uint64_t seed, uint32_t stream, uint64_t i, double f[3]) {
    if (f) er::widom_point(seed, stream, i, f);
}

void er_henry_result(double sum_b, double sum_eb, double n_total, double temperature,
                     double density, double *kh_mol_per_kg_pa, double *kh_mol_per_m3_pa,
                     double *e_avg_k, double *e_avg_kj_per_mol) {
    // src/henry.cu:166-174
    const double e_avg = sum_eb / sum_b;
    const double kh = sum_b / n_total / temperature / 8.314;
    if (e_avg_k) *e_avg_k = e_avg;
    if (e_avg_kj_per_mol) *e_avg_kj_per_mol = e_avg * 8.314 / 1000.0;
    if (kh_mol_per_m3_pa) *kh_mol_per_m3_pa = kh;
    if (kh_mol_per_kg_pa) *kh_mol_per_kg_pa = kh / density;
}

int er_batch_create(er_context *ctx, const er_task *tasks, int n_tasks, int use_graph,
                    er_batch **out) {
    if (!ctx || !out || (n_tasks > 0 && !tasks) || n_tasks < 0)
        return arg_error("er_batch_create: bad arguments");
    return guarded([&]() -> int {
        std::vector<er::Task> tk(n_tasks);
        for (int i = 0; i < n_tasks; i++) {
            tk[i].fw = F(tasks[i].fw);
            for (int d = 0; d < 3; d++) tk[i].grid_n[d] = tasks[i].grid_n[d];
            tk[i].precision = tasks[i].precision;
            tk[i].n_ins = tasks[i].n_insertions;
            tk[i].temperature = tasks[i].temperature;
            tk[i].seed = tasks[i].seed;
        }
        er::Batch *b = nullptr;
        int st = er::batch_create(C(ctx), tk.data(), n_tasks, use_graph != 0, &b);
        if (st == ST_OK) *out = reinterpret_cast<er_batch *>(b);
        return st;
    });
}

int er_batch_run(er_batch *batch) {
    if (!batch) return arg_error("er_batch_run: null batch");
    return guarded([&]() -> int { return er::batch_run(B(batch)); });
}

int er_batch_wait(er_batch *batch) {
    if (!batch) return arg_error("er_batch_wait: null batch");
    return guarded([&]() -> int { return er::batch_wait(B(batch)); });
}

int er_batch_result(er_batch *batch, int task, er_widom_result *out, er_grid **grid_out) {
    if (!batch) return arg_error("er_batch_result: null batch");
    er::WidomResult r;
    er::GridDev *g = nullptr;
    int st = er::batch_result(B(batch), task, &r, &g);
    if (st != ST_OK) return st;
    if (out) {
        out->sum_b = r.sum_b;
        out->sum_eb = r.sum_eb;
        out->n = r.n;
    }
    if (grid_out) *grid_out = reinterpret_cast<er_grid *>(g);
    return ST_OK;
}

int er_batch_set_timing(er_batch *batch, int on) {
    if (!batch) return arg_error("er_batch_set_timing: null batch");
    return guarded([&]() -> int { return er::batch_set_timing(B(batch), on != 0); });
}

int er_batch_kernel_ms(er_batch *batch, double *ms, int *launches) {
    if (!batch || !ms) return arg_error("er_batch_kernel_ms: null argument");
    return guarded([&]() -> int { return er::batch_kernel_ms(B(batch), ms, launches); });
}

int er_batch_release(er_batch *batch) {
    delete B(batch);
    return ST_OK;
}

int er_grid_upload(er_context *ctx, const float *host, const int n[3], er_grid **out) {
    if (!ctx || !host || !n || !out) return arg_error("er_grid_upload: null argument");
    return guarded([&]() -> int {
        er::GridDev *g = nullptr;
        int st = er::grid_upload(C(ctx), host, n, &g);
        if (st == ST_OK) *out = reinterpret_cast<er_grid *>(g);
        return st;
    });
}

int er_pockets_label(er_context *ctx, er_grid *grid, double temperature, int block_in_place,
                     int *n_pockets, int32_t *segments_out) {
    if (!ctx || !grid || !n_pockets) return arg_error("er_pockets_label: null argument");
    return guarded([&]() -> int {
        return er::pockets_label(C(ctx), G(grid), temperature, block_in_place != 0, n_pockets,
                                 segments_out);
    });
}

int er_pockets_label_batch(er_context *ctx, er_grid *const *grids, int n_grids,
                           double temperature, int block_in_place, int *n_pockets) {
    if (!ctx || (n_grids > 0 && (!grids || !n_pockets)) || n_grids < 0)
        return arg_error("er_pockets_label_batch: bad arguments");
    return guarded([&]() -> int {
        std::vector<er::GridDev *> g(n_grids);
        for (int i = 0; i < n_grids; i++) g[i] = G(grids[i]);
        return er::pockets_label_many(C(ctx), g.data(), n_grids, temperature,
                                      block_in_place != 0, n_pockets, nullptr);
    });
}

int er_grid_energies(er_context *ctx, const er_grid *grid, const double *frac_xyz,
                     uint64_t n_positions, double *energies_out) {
    if (!ctx || !grid) return arg_error("er_grid_energies: null argument");
    return guarded([&]() -> int {
        return er::grid_energies(C(ctx), const_cast<er::GridDev *>(G(grid)), frac_xyz,
                                 n_positions, energies_out);
    });
}

int er_widom_grid(er_context *ctx, er_grid *grid, uint64_t n_insertions, double temperature,
                  uint64_t seed, er_widom_result *out, double *energies_out) {
    if (!ctx || !grid || !out) return arg_error("er_widom_grid: null argument");
    return guarded([&]() -> int {
        er::WidomResult r;
        int st = er::widom_grid_run(C(ctx), G(grid), n_insertions, temperature, seed, &r,
                                    energies_out);
        if (st == ST_OK) {
            out->sum_b = r.sum_b;
            out->sum_eb = r.sum_eb;
            out->n = r.n;
        }
        return st;
    });
}

int er_server_create(er_context *ctx, const er_server_config *config, er_server **out) {
    if (!ctx || !out) return arg_error("er_server_create: null argument");
    return guarded([&]() -> int {
        er::ServerConfig cfg;
        if (config) {
            cfg.streams = config->streams;
            cfg.items_per_launch = config->items_per_launch;
            cfg.max_queue = config->max_queue;
            cfg.cache_bytes = config->cache_bytes;
            cfg.fail_launch = config->fail_launch;
        }
        if (cfg.streams < 1 || cfg.streams > 8 || cfg.items_per_launch < 0 || cfg.max_queue < 1 ||
            cfg.cache_bytes < -1 || cfg.fail_launch < 0)
            return arg_error("er_server_create: streams 1..8, items_per_launch >= 0, "
                             "max_queue >= 1, cache_bytes >= -1, fail_launch >= 0");
        std::unique_ptr<er::Server> srv(new er::Server(C(ctx), cfg));
        if (srv->status() != ST_OK) {
            er::set_error("er_server_create: no device memory for the launch slots");
            return srv->status();
        }
        *out = reinterpret_cast<er_server *>(srv.release());
        return ST_OK;
    });
}

int er_server_submit(er_server *server, const er_request *rq, uint64_t *ticket) {
    if (!server || !rq || !ticket || (rq->n_atoms > 0 && !rq->atoms) || rq->n_atoms < 0)
        return arg_error("er_server_submit: bad arguments");
    const bool grid = rq->grid_n[0] || rq->grid_n[1] || rq->grid_n[2];
    if (grid)
        for (int d = 0; d < 3; d++)
            if (rq->grid_n[d] < 2 || rq->grid_n[d] > 65536)
                return arg_error("er_server_submit: grid dimensions must be in [2, 65536]");
    if (grid && rq->precision != 0 && rq->precision != 1)
        return arg_error("er_server_submit: precision must be 0 or 1");
    if (rq->n_insertions && (rq->n_insertions >= (1ULL << 32) || !(rq->temperature > 0.0)))
        return arg_error("er_server_submit: insertions in [1, 2^32) and positive temperature");
    if (!grid && !rq->n_insertions) return arg_error("er_server_submit: empty request");
    return guarded([&]() -> int {
        er::RequestSpec spec;
        spec.atoms.assign(reinterpret_cast<const er::Particle *>(rq->atoms),
                          reinterpret_cast<const er::Particle *>(rq->atoms) + rq->n_atoms);
        std::memcpy(spec.t, rq->t_matrix, sizeof(spec.t));
        std::memcpy(spec.rep, rq->rep, sizeof(spec.rep));
        spec.rc2 = rq->rc2;
        std::memcpy(spec.grid_n, rq->grid_n, sizeof(spec.grid_n));
        spec.precision = rq->precision;
        spec.n_ins = rq->n_insertions;
        spec.temperature = rq->temperature;
        spec.seed = rq->seed;
        return reinterpret_cast<er::Server *>(server)->submit(spec, ticket);
    });
}

int er_server_poll(er_server *server, uint64_t ticket, double wait_ms, er_request_status *out) {
    if (!server || !out) return arg_error("er_server_poll: null argument");
    return guarded([&]() -> int {
        er::RequestInfo info;
        int st = reinterpret_cast<er::Server *>(server)->poll(ticket, &info, wait_ms);
        if (st != ST_OK) return st;
        out->state = info.state;
        out->status = info.status;
        out->t_submit = info.t_submit;
        out->t_admit = info.t_admit;
        out->t_done = info.t_done;
        out->widom.sum_b = info.widom.sum_b;
        out->widom.sum_eb = info.widom.sum_eb;
        out->widom.n = info.widom.n;
        return ST_OK;
    });
}

int er_server_fetch_grid(er_server *server, uint64_t ticket, void *host_out) {
    if (!server || !host_out) return arg_error("er_server_fetch_grid: null argument");
    return guarded([&]() -> int {
        return reinterpret_cast<er::Server *>(server)->fetch_grid(ticket, host_out);
    });
}

int er_server_release(er_server *server, uint64_t ticket) {
    if (!server) return arg_error("er_server_release: null server");
    return guarded([&]() -> int { return reinterpret_cast<er::Server *>(server)->release(ticket); });
}

int er_server_stats(er_server *server, int64_t *submitted, int64_t *rejected, int64_t *completed,
                    int64_t *launches) {
    if (!server) return arg_error("er_server_stats: null server");
    long long a = 0, b = 0, c = 0, d = 0;
    reinterpret_cast<er::Server *>(server)->stats(&a, &b, &c, &d);
    if (submitted) *submitted = a;
    if (rejected) *rejected = b;
    if (completed) *completed = c;
    if (launches) *launches = d;
    return ST_OK;
}

int er_server_cache_stats(er_server *server, int64_t *hits, int64_t *misses, int64_t *evictions,
                          int64_t *resident_frameworks, int64_t *resident_bytes) {
    if (!server) return arg_error("er_server_cache_stats: null server");
    return guarded([&]() -> int {
        const er::CacheStats c = reinterpret_cast<er::Server *>(server)->cache_stats();
        if (hits) *hits = c.hits;
        if (misses) *misses = c.misses;
        if (evictions) *evictions = c.evictions;
        if (resident_frameworks) *resident_frameworks = c.resident_frameworks;
        if (resident_bytes) *resident_bytes = c.resident_bytes;
        return ST_OK;
    });
}

int er_server_destroy(er_server *server) {
    delete reinterpret_cast<er::Server *>(server);
    return ST_OK;
}

int er_host_setup(const char *root, const char *structure, const char *adsorbate, int mode,
                  er_host_run **out) {
    if (!structure || !adsorbate || !out || (mode != 0 && mode != 1))
        return arg_error("er_host_setup: bad arguments");
    return guarded([&]() -> int {
        er_host_run *r = new er_host_run();
        r->mode = mode;
        std::string msg;
        const std::string rt = root ? root : ".";
        int st = mode == 0 ? er::host::setup_grid_run(rt, structure, adsorbate, &r->gp, &r->setup,
                                                      &msg)
                           : er::host::setup_henry_run(rt, structure, adsorbate, &r->hp, &r->setup,
                                                       &msg);
        if (st != ST_OK) {
            er::set_error(msg);
            delete r;
            return st;
        }
        *out = r;
        return ST_OK;
    });
}

int er_host_run_info(const er_host_run *r, er_host_info *o) {
    if (!r || !o) return arg_error("er_host_run_info: null argument");
    std::memset(o, 0, sizeof(*o));
    const er::host::Framework &fw = r->setup.fw;
    o->a = fw.a;
    o->b = fw.b;
    o->c = fw.c;
    o->alpha = fw.alpha;
    o->beta = fw.beta;
    o->gamma = fw.gamma;
    o->n_atoms = fw.noatoms;
    o->density = fw.density;
    o->volume = fw.volume_unitcell;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            o->t_matrix[3 * i + j] = fw.t_matrix[i][j];
            o->inv_t_matrix[3 * i + j] = fw.inv_t_matrix[i][j];
        }
    if (r->mode == 0) {
        const er::host::GridParams &p = r->gp;
        for (int d = 0; d < 3; d++) o->rep[d] = p.rep[d];
        o->r_cutoff_squared = p.r_cutoff_squared;
        o->temperature = p.T;
        o->epsilon_guest = p.epsilon_guest;
        o->sigma_guest = p.sigma_guest;
        o->adsorbate_mw = p.adsorbateMW;
        o->grid_n[0] = p.N_x;
        o->grid_n[1] = p.N_y;
        o->grid_n[2] = p.N_z;
        o->grid_resolution = p.grid_resolution;
        o->grid_precision = p.precision;
    } else {
        const er::host::HenryParams &p = r->hp;
        for (int d = 0; d < 3; d++) o->rep[d] = p.rep[d];
        o->r_cutoff_squared = p.r_cutoff_squared;
        o->temperature = p.T;
        o->epsilon_guest = p.epsilon_guest;
        o->sigma_guest = p.sigma_guest;
        o->adsorbate_mw = p.adsorbateMW;
        o->insertions_per_a3 = p.numinsertionsperA3;
        o->numinsertions = p.numinsertions;
        o->insertions_per_thread = p.numinsertionsperthread;
        o->total_insertions = er::host::henry_total(p);
        o->seed = p.seed;
    }
    return ST_OK;
}

int er_host_run_particles(const er_host_run *r, er_particle *out, int capacity) {
    if (!r || !out) return arg_error("er_host_run_particles: null argument");
    const int n = (int)r->setup.atoms.size();
    if (capacity < n) return arg_error("er_host_run_particles: capacity too small");
    std::memcpy(out, r->setup.atoms.data(), sizeof(er_particle) * n);
    return ST_OK;
}

int er_host_run_label(const er_host_run *r, int i, char *out, int capacity) {
    if (!r || !out || capacity < 1 || i < 0 || i >= r->setup.fw.noatoms)
        return arg_error("er_host_run_label: bad arguments");
    std::snprintf(out, (size_t)capacity, "%s", r->setup.fw.identity[i].c_str());
    return ST_OK;
}

int er_host_run_write_settings(const er_host_run *r, const char *path) {
    if (!r || !path) return arg_error("er_host_run_write_settings: null argument");
    FILE *f = std::fopen(path, "w");
    if (!f) {
        er::set_error(std::string("cannot open ") + path);
        return ST_IO;
    }
    if (r->mode == 0)
        er::host::write_grid_settings(f, r->gp, r->setup.fw, r->setup.ff, r->setup.atoms.data());
    else
        er::host::write_henry_settings(f, r->hp, r->setup.fw, r->setup.ff, r->setup.atoms.data());
    std::fclose(f);
    return ST_OK;
}

void er_host_run_release(er_host_run *r) { delete r; }

int er_write_grid_file(const char *path, const char *format, const double *e64, const float *e32,
                       const int n[3], const double t_matrix[9], int threads) {
    if (!path || !format || !n || (!e64 == !e32) || !t_matrix)
        return arg_error("er_write_grid_file: bad arguments");
    const std::string fm = format;
    if (fm != "txt" && fm != "cube") return arg_error("Grid output format must be txt or cube\n");
    FILE *f = std::fopen(path, "w");
    if (!f) {
        er::set_error(std::string("cannot open ") + path);
        return ST_IO;
    }
    double t[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[i][j] = t_matrix[3 * i + j];
    if (fm == "txt") er::host::write_txt_header(f, n);
    else er::host::write_cube_header(f, n, t);
    int st = er::host::write_grid_values(f, fm, e64, e32, n, threads);
    std::fclose(f);
    if (st != ST_OK) er::set_error("er_write_grid_file: write failed");
    return st;
}

int er_accessibility_grid(er_context *ctx, const er_particle *atoms, int n_atoms,
                          const double t_matrix[9], const int n[3], double *out) {
    if (!ctx || !t_matrix || !n || !out || n_atoms < 0 || (n_atoms > 0 && !atoms))
        return arg_error("er_accessibility_grid: bad arguments");
    return guarded([&]() -> int {
        return er::accessibility_grid(C(ctx), reinterpret_cast<const er::Particle *>(atoms),
                                      n_atoms, t_matrix, n, out);
    });
}

int er_direct_grid(er_context *ctx, const er_particle *atoms, int n_atoms,
                   const double t_matrix[9], const int rep[3], double rc2, const int n[3],
                   double *host_out, er_grid **resident_out) {
    if (!ctx || !t_matrix || !rep || !n || n_atoms < 0 || (n_atoms > 0 && !atoms))
        return arg_error("er_direct_grid: bad arguments");
    return guarded([&]() -> int {
        er::GridDev *g = nullptr;
        int st = er::direct_grid(C(ctx), reinterpret_cast<const er::Particle *>(atoms), n_atoms,
                                 t_matrix, rep, rc2, n, host_out, resident_out ? &g : nullptr);
        if (st == ST_OK && resident_out) *resident_out = reinterpret_cast<er_grid *>(g);
        return st;
    });
}

int er_direct_widom(er_context *ctx, const er_particle *atoms, int n_atoms,
                    const double t_matrix[9], const int rep[3], double rc2,
                    uint64_t n_insertions, double temperature, uint64_t seed,
                    er_widom_result *out, double *energies_out) {
    if (!ctx || !t_matrix || !rep || !out || n_atoms < 0 || (n_atoms > 0 && !atoms))
        return arg_error("er_direct_widom: bad arguments");
    return guarded([&]() -> int {
        er::WidomResult r;
        int st = er::direct_widom(C(ctx), reinterpret_cast<const er::Particle *>(atoms), n_atoms,
                                  t_matrix, rep, rc2, n_insertions, temperature, seed, &r,
                                  energies_out);
        if (st == ST_OK) {
            out->sum_b = r.sum_b;
            out->sum_eb = r.sum_eb;
            out->n = r.n;
        }
        return st;
    });
}

int er_grid_file_info(const char *path, int n[3]) {
    if (!path || !n) return arg_error("er_grid_file_info: null argument");
    return guarded([&]() -> int {
        std::string msg;
        std::vector<double> v;
        int st = er::host::read_grid_file(path, n, &v, 16, &msg);
        if (st) er::set_error(msg);
        return st;
    });
}

int er_read_grid_file(const char *path, const int n_expected[3], double *out) {
    if (!path || !n_expected || !out) return arg_error("er_read_grid_file: null argument");
    return guarded([&]() -> int {
        std::string msg;
        std::vector<double> v;
        int n[3];
        int st = er::host::read_grid_file(path, n, &v, 16, &msg);
        if (st) {
            er::set_error(msg);
            return st;
        }
        if (n[0] != n_expected[0] || n[1] != n_expected[1] || n[2] != n_expected[2])
            return arg_error("er_read_grid_file: grid dimensions differ from n_expected");
        std::memcpy(out, v.data(), sizeof(double) * v.size());
        return ST_OK;
    });
}

int er_write_grid_binary(const char *path, const double *e64, const float *e32, const int n[3],
                         const double t_matrix[9]) {
    if (!path || !n || !t_matrix || (!e64 == !e32))
        return arg_error("er_write_grid_binary: bad arguments");
    std::string msg;
    int st = er::host::write_grid_bin(path, n, t_matrix, e64, e32, &msg);
    if (st) er::set_error(msg);
    return st;
}

int er_grid_load(er_context *ctx, const char *path, er_grid **out) {
    if (!ctx || !path || !out) return arg_error("er_grid_load: null argument");
    return guarded([&]() -> int {
        std::string msg;
        std::vector<double> v;
        int n[3];
        int st = er::host::read_grid_file(path, n, &v, 16, &msg);
        if (st) {
            er::set_error(msg);
            return st;
        }
        std::vector<float> f(v.begin(), v.end());
        er::GridDev *g = nullptr;
        st = er::grid_upload(C(ctx), f.data(), n, &g);
        if (st == ST_OK) *out = reinterpret_cast<er_grid *>(g);
        return st;
    });
}

int er_replication_factors(const double t_matrix[9], double cutoff, int twice, int rep[3]) {
    if (!t_matrix || !rep) return arg_error("er_replication_factors: null argument");
    int st = er::host::replication_factors(t_matrix, cutoff, twice != 0, rep);
    if (st != ST_OK) er::set_error("er_replication_factors: degenerate cell or negative cutoff");
    return st;
}

int er_write_uc_file(const char *root, const char *structure, double cutoff, int twice,
                     int rep[3]) {
    if (!root || !structure) return arg_error("er_write_uc_file: null argument");
    return guarded([&]() -> int {
        std::string msg;
        double t[9];
        int r[3];
        int st = er::host::read_cssr_cell(root, structure, t, &msg);
        if (st == ST_OK) {
            st = er::host::replication_factors(t, cutoff, twice != 0, r);
            if (st != ST_OK) msg = "degenerate cell or negative cutoff";
        }
        if (st == ST_OK) st = er::host::write_uc_file(root, structure, twice != 0, r, &msg);
        if (st != ST_OK) {
            er::set_error(msg);
            return st;
        }
        if (rep) std::memcpy(rep, r, sizeof(r));
        return ST_OK;
    });
}

}  // extern "C"
