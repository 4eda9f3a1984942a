// er_runtime.cpp -- see er_runtime.h.
#include "er_runtime.h"

#include <algorithm>
#include <memory>
#include <cmath>
#include <cstring>
#include <sstream>

namespace er {

namespace {
thread_local std::string g_last_error;

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::ostringstream os_;                                                        \
            os_ << "HIP error " << hipGetErrorName(e_) << " (" << hipGetErrorString(e_)    \
                << ") at " << __FILE__ << ":" << __LINE__ << " in " << #expr;             \
            set_error(os_.str());                                                          \
            return ST_HIP;                                                                 \
        }                                                                                  \
    } while (0)

void invert3(const double t[9], double inv[9]) {
    const double c00 = t[4] * t[8] - t[5] * t[7];
    const double c01 = t[5] * t[6] - t[3] * t[8];
    const double c02 = t[3] * t[7] - t[4] * t[6];
    const double det = t[0] * c00 + t[1] * c01 + t[2] * c02;
    inv[0] = c00 / det;
    inv[1] = (t[2] * t[7] - t[1] * t[8]) / det;
    inv[2] = (t[1] * t[5] - t[2] * t[4]) / det;
    inv[3] = c01 / det;
    inv[4] = (t[0] * t[8] - t[2] * t[6]) / det;
    inv[5] = (t[2] * t[3] - t[0] * t[5]) / det;
    inv[6] = c02 / det;
    inv[7] = (t[1] * t[6] - t[0] * t[7]) / det;
    inv[8] = (t[0] * t[4] - t[1] * t[3]) / det;
}

double det3(const double t[9]) {
    return t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
           t[2] * (t[3] * t[7] - t[4] * t[6]);
}

// Perpendicular width of the cell along fractional axis d: 1/|row_d(T^-1)|.
double perp_width(const double inv[9], int d) {
    const double *r = inv + 3 * d;
    return 1.0 / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
}

template <class T>
int pool_alloc(Context *ctx, size_t count, T **out, std::vector<void *> *owned) {
    void *p = ctx->pool.alloc(count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (!p) return ST_NOMEM;
    *out = (T *)p;
    if (owned) owned->push_back(p);
    return ST_OK;
}

#define ST_TRY(expr)                    \
    do {                                \
        int s_ = (expr);                \
        if (s_ != ST_OK) return s_;     \
    } while (0)

}  // namespace

void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Device pool
// ---------------------------------------------------------------------------

size_t DevicePool::round_size(size_t b) {
    if (b <= (1u << 20)) {
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    const size_t mb = 1u << 20;
    return (b + mb - 1) / mb * mb;
}

void *DevicePool::alloc(size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    const size_t sz = round_size(bytes);
    auto it = cache_.find(sz);
    if (it != cache_.end() && !it->second.empty()) {
        void *p = it->second.back();
        it->second.pop_back();
        live_[p] = sz;
        in_use_ += sz;
        return p;
    }
    if (ctx_->hbm_cap && reserved_ + sz > ctx_->hbm_cap) {
        // give cached blocks back before refusing
        for (auto &kv : cache_)
            for (void *p : kv.second) {
                (void)hipFree(p);
                reserved_ -= kv.first;
            }
        cache_.clear();
        if (reserved_ + sz > ctx_->hbm_cap) {
            std::ostringstream os;
            os << "HBM cap exceeded: need " << sz << " B with " << reserved_ << " of "
               << ctx_->hbm_cap << " B reserved";
            set_error(os.str());
            return nullptr;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        std::ostringstream os;
        os << "hipMalloc of " << sz << " B failed";
        set_error(os.str());
        return nullptr;
    }
    live_[p] = sz;
    in_use_ += sz;
    reserved_ += sz;
    return p;
}

void DevicePool::free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = live_.find(p);
    if (it == live_.end()) return;
    cache_[it->second].push_back(p);
    in_use_ -= it->second;
    live_.erase(it);
}

void DevicePool::trim() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : cache_)
        for (void *p : kv.second) {
            (void)hipFree(p);
            reserved_ -= kv.first;
        }
    cache_.clear();
}

DevicePool::~DevicePool() {
    for (auto &kv : cache_)
        for (void *p : kv.second) (void)hipFree(p);
    for (auto &kv : live_) (void)hipFree(kv.first);
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

Context::~Context() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    pool.free(d_counter);
    pool.free(d_partials);
    pool.free(d_sums);
    pool.free(d_scratch);
    if (h_stage) (void)hipHostFree(h_stage);
}

void *Context::host_stage(size_t bytes) {
    if (bytes <= h_stage_bytes) return h_stage;
    if (h_stage) {
        (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
    }
    const size_t b = std::max<size_t>(bytes, 1 << 20);
    if (hipHostMalloc(&h_stage, b) != hipSuccess) {
        (void)hipGetLastError();
        h_stage = nullptr;
        set_error("pinned host staging allocation failed");
        return nullptr;
    }
    h_stage_bytes = b;
    return h_stage;
}

int context_create(int device, size_t hbm_cap, Context **out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return ST_NODEVICE;
    }
    if (device < 0 || device >= ndev) {
        set_error("device index out of range");
        return ST_ARG;
    }
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    Context *c = new Context();
    c->device = device;
    c->hbm_cap = hbm_cap;
    c->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreateWithFlags failed");
        return ST_HIP;
    }
    if (pool_alloc(c, 1, &c->d_counter, nullptr) != ST_OK ||
        pool_alloc(c, kReduceBlocks, &c->d_partials, nullptr) != ST_OK ||
        pool_alloc(c, 2, &c->d_sums, nullptr) != ST_OK ||
        scratch_ensure(c, c->stream, &c->d_scratch, &c->scratch_wgs) != ST_OK) {
        delete c;
        return ST_NOMEM;
    }
    *out = c;
    return ST_OK;
}

int tiles_workgroups(const Context *ctx, int math, long long items, int scratch_wgs) {
    const int occ = tiles_occupancy(math, ctx->tuning.ppl, ctx->tuning.cap);
    long long w = (long long)ctx->n_cu * occ;
    if (w > items) w = items;
    if (w > scratch_wgs) w = scratch_wgs;
    return (int)std::max<long long>(w, 1);
}

int scratch_ensure(Context *ctx, hipStream_t s, uint32_t **buf, int *wgs) {
    int occ = 1;
    for (int m : {(int)MATH_MIXED, (int)MATH_FP64, (int)MATH_COUNT, (int)MATH_OVERLAP})
        occ = std::max(occ, tiles_occupancy(m, ctx->tuning.ppl, ctx->tuning.cap));
    int need = ctx->n_cu * occ;
    if (*buf && *wgs >= need) return ST_OK;
    // Under a tight HBM cap take fewer slices: launches then use fewer
    // workgroups, which changes no result.
    void *p = nullptr;
    for (; need >= 1 && need > *wgs; need /= 2) {
        p = ctx->pool.alloc((size_t)need * kRegionRecords * sizeof(uint32_t));
        if (p) break;
    }
    if (!p) {
        if (*buf) return ST_OK;  // keep the smaller scratch
        set_error("no device memory for the region scratch");
        return ST_NOMEM;
    }
    if (*buf) {
        HIP_TRY(hipStreamSynchronize(s));
        ctx->pool.free(*buf);
    }
    *buf = (uint32_t *)p;
    *wgs = need;
    return ST_OK;
}

// ---------------------------------------------------------------------------
// Framework upload
// ---------------------------------------------------------------------------

FrameworkDev::~FrameworkDev() {
    if (ctx)
        for (void *p : owned) ctx->pool.free(p);
}

int framework_upload(Context *ctx, const Particle *atoms, int n, const double t[9],
                     const int rep[3], double rc2, FrameworkDev **out) {
    if (!ctx || (!atoms && n > 0) || n < 0 || !t || !rep || !out) {
        set_error("framework_upload: null argument or negative atom count");
        return ST_ARG;
    }
    if (rep[0] < 0 || rep[1] < 0 || rep[2] < 0 || rep[0] > 64 || rep[1] > 64 || rep[2] > 64) {
        set_error("framework_upload: replication factors must be in [0, 64]");
        return ST_ARG;
    }
    if (!(rc2 > 0.0) || !std::isfinite(rc2)) {
        set_error("framework_upload: rc2 must be positive and finite");
        return ST_ARG;
    }
    const double det = det3(t);
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) {
        set_error("framework_upload: singular cell matrix");
        return ST_ARG;
    }
    for (int i = 0; i < n; i++) {
        const Particle &p = atoms[i];
        if (!std::isfinite(p.x_f) || !std::isfinite(p.y_f) || !std::isfinite(p.z_f) ||
            !std::isfinite(p.eps) || !std::isfinite(p.sig2) || std::fabs(p.x_f) > 1e6 ||
            std::fabs(p.y_f) > 1e6 || std::fabs(p.z_f) > 1e6) {
            set_error("framework_upload: non-finite or out-of-range atom");
            return ST_ARG;
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    FrameworkDev *fw = new FrameworkDev();
    fw->ctx = ctx;
    fw->n_atoms = n;
    std::memcpy(fw->t, t, sizeof(fw->t));
    invert3(t, fw->inv_t);
    fw->rc2 = rc2;
    fw->rc = std::sqrt(rc2);
    fw->volume = std::fabs(det);
    fw->s2_switch = ctx->tuning.s2_switch;
    std::memcpy(fw->rep, rep, sizeof(fw->rep));
    fw->n_images = (long long)(2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1);

    // Pair parameters: one type per distinct (eps_ij, sig2_ij), first-seen order.
    std::vector<int> type_of(n);
    std::map<std::pair<uint64_t, uint64_t>, int> seen;
    for (int i = 0; i < n; i++) {
        uint64_t ke, ks;
        std::memcpy(&ke, &atoms[i].eps, 8);
        std::memcpy(&ks, &atoms[i].sig2, 8);
        auto key = std::make_pair(ke, ks);
        auto it = seen.find(key);
        if (it == seen.end()) {
            int id = (int)fw->types.size();
            seen[key] = id;
            TypeParams tp;
            tp.sig2 = atoms[i].sig2;
            tp.eps4 = 4.0 * atoms[i].eps;
            tp.rclose = std::sqrt(std::max(atoms[i].sig2, 0.0) / ctx->tuning.s2_switch);
            fw->types.push_back(tp);
            type_of[i] = id;
        } else {
            type_of[i] = it->second;
        }
    }
    const int nt = (int)fw->types.size();
    fw->ntypes = nt;
    std::vector<int> type_atom0(nt + 1, 0), type_block0(nt + 1, 0);
    for (int i = 0; i < n; i++) type_atom0[type_of[i] + 1]++;
    for (int k = 0; k < nt; k++) type_atom0[k + 1] += type_atom0[k];
    std::vector<AtomDev> grouped(n > 0 ? n : 1);
    {
        std::vector<int> fill(nt, 0);
        for (int i = 0; i < n; i++) {
            int k = type_of[i];
            AtomDev &a = grouped[type_atom0[k] + fill[k]++];
            a.fx = atoms[i].x_f;
            a.fy = atoms[i].y_f;
            a.fz = atoms[i].z_f;
            a.type = k;
            a.pad = 0;
        }
    }
    for (int k = 0; k < nt; k++) {
        long long work = (long long)(type_atom0[k + 1] - type_atom0[k]) * fw->n_images;
        long long blocks = (work + kWG - 1) / kWG;
        if (type_block0[k] + blocks > (1LL << 30)) {
            delete fw;
            set_error("framework_upload: too many image atoms");
            return ST_ARG;
        }
        type_block0[k + 1] = type_block0[k] + (int)blocks;
    }
    NearPlan plan;
    plan.ntypes = nt;
    std::memcpy(plan.rep, rep, sizeof(plan.rep));
    std::memcpy(plan.t, t, sizeof(plan.t));
    size_t cap_atoms = 1;
    for (int d = 0; d < 3; d++) {
        const double row = 1.0 / perp_width(fw->inv_t, d);
        plan.ext[d] = fw->rc * row + 1e-7;
        const long long per_axis =
            std::min<long long>(2LL * rep[d] + 1, (long long)std::floor(1.0 + 2.0 * plan.ext[d]) + 1);
        cap_atoms *= (size_t)std::max<long long>(per_axis, 1);
    }
    plan.nblocks = type_block0[nt];
    plan.n_images = (int)fw->n_images;
    fw->near_capacity = std::max<size_t>(1, (size_t)n * cap_atoms);
    int st = ST_OK;
    if ((st = pool_alloc(ctx, n > 0 ? n : 1, &fw->d_atoms, &fw->owned)) ||
        (st = pool_alloc(ctx, nt + 1, &fw->d_type_atom0, &fw->owned)) ||
        (st = pool_alloc(ctx, nt + 1, &fw->d_type_block0, &fw->owned)) ||
        (st = pool_alloc(ctx, nt + 1, &fw->d_type_begin, &fw->owned)) ||
        (st = pool_alloc(ctx, nt > 0 ? nt : 1, &fw->d_types, &fw->owned)) ||
        (st = pool_alloc(ctx, fw->near_capacity, &fw->d_near, &fw->owned)) ||
        (st = pool_alloc(ctx, plan.nblocks + 1, &fw->d_block_count, &fw->owned)) ||
        (st = pool_alloc(ctx, plan.nblocks + 1, &fw->d_block_off, &fw->owned))) {
        delete fw;
        return st;
    }
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(fw->d_atoms, grouped.data(), sizeof(AtomDev) * std::max(n, 1),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(fw->d_type_atom0, type_atom0.data(), sizeof(int) * (nt + 1),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(fw->d_type_block0, type_block0.data(), sizeof(int) * (nt + 1),
                           hipMemcpyHostToDevice, s));
    if (nt > 0)
        HIP_TRY(hipMemcpyAsync(fw->d_types, fw->types.data(), sizeof(TypeParams) * nt,
                               hipMemcpyHostToDevice, s));
    plan.atoms = fw->d_atoms;
    plan.type_atom0 = fw->d_type_atom0;
    plan.type_block0 = fw->d_type_block0;
    HIP_TRY(launch_near_count(plan, fw->d_block_count, s));
    HIP_TRY(launch_near_offsets(plan, fw->d_block_count, fw->d_block_off, fw->d_type_begin, s));
    HIP_TRY(launch_near_write(plan, fw->d_block_off, fw->d_near, s));
    HIP_TRY(hipStreamSynchronize(s));
    *out = fw;
    return ST_OK;
}

// ---------------------------------------------------------------------------
// Job construction
// ---------------------------------------------------------------------------

int grid_size_class(const double t[9], const int n[3], const Tuning &tu) {
    double lg = 0.0;
    for (int d = 0; d < 3; d++) {
        const double a = std::sqrt(t[d] * t[d] + t[3 + d] * t[3 + d] + t[6 + d] * t[6 + d]);
        lg += std::log(a / std::max(n[d] - 1, 1));
    }
    const double h = std::exp(lg / 3.0);
    if (h < tu.class_spacing[0]) return 0;
    if (h < tu.class_spacing[1]) return 1;
    return 2;
}

void grid_tile_shape(const double t[9], const int n[3], const Tuning &tu, int tile[3]) {
    const int *ct = tu.class_tile[grid_size_class(t, n, tu)];
    const bool set = ct[0] > 0 && ct[1] > 0 && ct[2] > 0;
    for (int d = 0; d < 3; d++) tile[d] = set ? ct[d] : tu.tile[d];
}

void choose_grid_tile(const FrameworkDev *fw, const int n[3], const Tuning &tu, int tile[3]) {
    grid_tile_shape(fw->t, n, tu, tile);
}

// Bins of near-cubic shape (per perpendicular width) whose mean occupancy is
// at most widom_per_bin: the tile kernel takes a bin's insertions 256 x
// (points per lane, at most 2) at a time, so a mean just below a multiple of
// that keeps the last chunk full while bins above the mean stay rare.  Refines the rounded counts: adds bins
// along the axis with the widest bins until the mean fits, then removes any
// bin whose removal keeps it fitting.
void choose_widom_bins(const FrameworkDev *fw, uint64_t n_ins, const Tuning &tu, int nb[3]) {
    double w[3];
    for (int d = 0; d < 3; d++) w[d] = perp_width(fw->inv_t, d);
    const double per = (double)std::max(tu.widom_per_bin, 1);
    double target = (double)n_ins / per;
    target = std::min(std::max(target, 1.0), (double)(1 << 21));
    const double k = std::cbrt(target / (w[0] * w[1] * w[2]));
    for (int d = 0; d < 3; d++) nb[d] = std::max(1, std::min(1 << 10, (int)std::lround(w[d] * k)));
    auto mean = [&]() { return (double)n_ins / ((double)nb[0] * nb[1] * nb[2]); };
    while (mean() > per) {
        int d = -1;
        for (int e = 0; e < 3; e++)
            if (nb[e] < (1 << 10) && (d < 0 || w[e] / nb[e] > w[d] / nb[d])) d = e;
        if (d < 0) break;
        nb[d]++;
    }
    for (bool changed = true; changed;) {
        changed = false;
        int d = -1;  // narrowest bins first
        for (int e = 0; e < 3; e++)
            if (nb[e] > 1 && (d < 0 || w[e] / nb[e] < w[d] / nb[d])) d = e;
        if (d < 0) break;
        nb[d]--;
        if (mean() <= per) changed = true;
        else nb[d]++;
    }
}

// Cells (grid tiles or Widom bins) per region per axis: about region_len
// Angstrom along each lattice vector, shrunk (largest extent first) until the
// expected sub-list -- the framework's atom density times the sphere of
// radius r_c + region radius -- fills at most 3/4 of a workgroup's scratch
// and the job has at least min_items regions (so a job launched alone still
// fills the device).  cell_frac[d] is a cell's fractional extent, ncell[d]
// the cells per axis.  Regions of one cell skip the sub-list.  A cell's
// candidate list, and so every result, is the same whatever region it is in.
void choose_region(const FrameworkDev *fw, const double cell_frac[3], const int ncell[3],
                   const Tuning &tu, long long min_items, int region[3], int slice_records) {
    double edge[3];
    for (int d = 0; d < 3; d++) {
        const double a = std::sqrt(fw->t[d] * fw->t[d] + fw->t[3 + d] * fw->t[3 + d] +
                                   fw->t[6 + d] * fw->t[6 + d]);
        edge[d] = a * cell_frac[d];
        region[d] = 1;
        if (tu.region_len > 0.0 && edge[d] > 0.0)
            region[d] = std::max(1, (int)std::lround(tu.region_len / edge[d]));
        region[d] = std::min(region[d], 4096);
    }
    const double pi = 3.14159265358979323846;
    const double dens = fw->n_atoms / std::max(fw->volume, 1e-9);
    for (;;) {
        // radius of the region box: largest centre-to-corner distance
        double r2 = 0.0;
        for (int c = 0; c < 8; c++) {
            double f[3];
            for (int d = 0; d < 3; d++)
                f[d] = ((c >> d) & 1 ? 0.5 : -0.5) * region[d] * cell_frac[d];
            double x = 0, y = 0, z = 0;
            x = fw->t[0] * f[0] + fw->t[1] * f[1] + fw->t[2] * f[2];
            y = fw->t[3] * f[0] + fw->t[4] * f[1] + fw->t[5] * f[2];
            z = fw->t[6] * f[0] + fw->t[7] * f[1] + fw->t[8] * f[2];
            r2 = std::max(r2, x * x + y * y + z * z);
        }
        const double rr = fw->rc + std::sqrt(r2) + 0.01;
        const double expect = dens * (4.0 / 3.0) * pi * rr * rr * rr;
        long long items = 1;
        for (int d = 0; d < 3; d++) items *= (ncell[d] + region[d] - 1) / region[d];
        if (expect <= 0.75 * slice_records && items >= min_items) break;
        int big = 0;
        for (int d = 1; d < 3; d++)
            if (region[d] * edge[d] > region[big] * edge[big]) big = d;
        if (region[big] == 1) break;
        region[big] = (region[big] + 1) / 2;
    }
}

void choose_grid_region(const FrameworkDev *fw, const int n[3], const int tile[3],
                        const Tuning &tu, int region[3], int slice_records) {
    double frac[3];
    int nt[3];
    for (int d = 0; d < 3; d++) {
        frac[d] = (double)tile[d] / (n[d] - 1);
        nt[d] = (n[d] + tile[d] - 1) / tile[d];
    }
    choose_region(fw, frac, nt, tu, kMinGridItems, region, slice_records);
}

JobDev make_grid_job(const FrameworkDev *fw, const int n[3], int precision, void *d_out,
                     const Tuning &tu, int math, int slice_records) {
    JobDev j;
    std::memset(&j, 0, sizeof(j));
    std::memcpy(j.t, fw->t, sizeof(j.t));
    j.rc2 = fw->rc2;
    j.rc = fw->rc;
    j.min_r2 = 1e-12 * 1e-12;  // src/computegridsheet.h:19,77
    j.near = fw->d_near;
    j.type_begin = fw->d_type_begin;
    j.types = fw->d_types;
    j.ntypes = fw->ntypes;
    j.kind = JOB_GRID;
    j.mode = math == MATH_COUNT ? OUT_COUNT : (precision == 1 ? OUT_F64 : OUT_F32);
    int tile[3], region[3];
    choose_grid_tile(fw, n, tu, tile);
    choose_grid_region(fw, n, tile, tu, region, slice_records);
    long long items = 1;
    for (int d = 0; d < 3; d++) {
        j.n[d] = n[d];
        j.tile[d] = tile[d];
        j.ntiles[d] = (n[d] + tile[d] - 1) / tile[d];
        j.region[d] = std::min(region[d], j.ntiles[d]);
        j.nregions[d] = (j.ntiles[d] + j.region[d] - 1) / j.region[d];
        items *= j.nregions[d];
        // point boxes; the quadrant of a tile holds its i range and half its
        // j and k ranges (used only where the kernel takes that layout)
        const int wave_pts = d == 0 ? tile[d] : std::max(tile[d] / 2, 1);
        j.cell_box[d] = 0.5 * (double)(tile[d] - 1) / (double)(n[d] - 1);
        j.wave_box[d] = 0.5 * (double)(wave_pts - 1) / (double)(n[d] - 1);
    }
    j.out = d_out;
    j.n_items = items;
    return j;
}

JobDev make_widom_job(const FrameworkDev *fw, uint64_t n_ins, uint64_t seed,
                      const WidomBuffers &wb, const Tuning &tu, long long min_items,
                      double temperature) {
    JobDev j;
    std::memset(&j, 0, sizeof(j));
    std::memcpy(j.t, fw->t, sizeof(j.t));
    j.rc2 = fw->rc2;
    j.rc = fw->rc;
    j.min_r2 = 1e-16 * 1e-16;  // src/doHenry.h:15,91
    j.near = fw->d_near;
    j.type_begin = fw->d_type_begin;
    j.types = fw->d_types;
    j.ntypes = fw->ntypes;
    j.kind = JOB_WIDOM;
    j.mode = OUT_F64;
    for (int d = 0; d < 3; d++) {
        j.nbins[d] = wb.nbins[d];
        j.cell_box[d] = 0.5 / (double)wb.nbins[d];
    }
    j.sorted = wb.sorted;
    j.bin_start = wb.start;
    j.seed = seed;
    j.stream = 0;
    j.e_out = wb.e;
    // saturation bound: every pair energy is >= -eps_ij when no eps_ij is
    // negative (exp(-800) is exactly 0 in FP64)
    bool eps_ok = true;
    for (const TypeParams &tp : fw->types) eps_ok = eps_ok && tp.eps4 >= 0.0;
    j.sat_e = (temperature > 0.0 && std::isfinite(temperature) && eps_ok) ? 800.0 * temperature
                                                                         : 0.0;
    // A far or shell pair lies beyond r_FP64 = sigma / sqrt(s2) of its point,
    // u = (sigma / r)^2 <= s2: 4 eps (u^6 - u^3) decreases in u up to
    // u = 2^(-1/3) (the LJ minimum, -eps), so it is at least 4 eps (s2^6 - s2^3)
    // below that and -eps beyond; 0.01 eps of margin covers the FP32 pair math.
    const double s2 = fw->s2_switch;
    const double fmin = (s2 > 0.0 && s2 < 0.7937005259840998)
                            ? 4.0 * (s2 * s2 * s2 * s2 * s2 * s2 - s2 * s2 * s2)
                            : -1.0;
    j.fs_floor = std::min(1.0, -fmin + 0.01);
    // regions of bins share one candidate sub-list, like regions of grid tiles
    double frac[3];
    for (int d = 0; d < 3; d++) frac[d] = 1.0 / wb.nbins[d];
    choose_region(fw, frac, wb.nbins, tu, min_items, j.region);
    long long items = 1;
    for (int d = 0; d < 3; d++) {
        j.region[d] = std::min(j.region[d], wb.nbins[d]);
        j.nregions[d] = (wb.nbins[d] + j.region[d] - 1) / j.region[d];
        items *= j.nregions[d];
    }
    j.n_items = items;
    (void)n_ins;
    return j;
}

int check_grid_dims(const int n[3]) {
    for (int d = 0; d < 3; d++)
        if (n[d] < 2 || n[d] > 65536) {
            set_error("grid dimensions must be in [2, 65536] (the reference divides by N-1)");
            return ST_ARG;
        }
    const long long p = (long long)n[0] * n[1] * n[2];
    if (p > (1LL << 36)) {
        set_error("grid too large");
        return ST_ARG;
    }
    return ST_OK;
}

namespace {

int persistent_wgs(Context *ctx, int math, long long items) {
    return tiles_workgroups(ctx, math, items, ctx->scratch_wgs);
}

// Launch one job array on the context stream with the context's counter.
int run_jobs(Context *ctx, int math, JobDev *d_jobs, int njobs, long long total,
             unsigned long long *d_counter) {
    hipStream_t s = ctx->stream;
    HIP_TRY(launch_fill_u64(d_counter, 0ull, 1, s));
    HIP_TRY(launch_tiles(math, ctx->tuning.ppl, d_jobs, njobs, total, d_counter,
                         persistent_wgs(ctx, math, total), ctx->tuning.cap, ctx->d_scratch,
                         kRegionRecords, s));
    return ST_OK;
}

struct Scoped {
    Context *ctx;
    std::vector<void *> ptrs;
    ~Scoped() {
        for (void *p : ptrs) ctx->pool.free(p);
    }
};

}  // namespace

int check_tuning(const Tuning &tu) {
    if (tu.ppl != 1 && tu.ppl != 2 && tu.ppl != 4) {
        set_error("tuning: points per lane must be 1, 2 or 4");
        return ST_ARG;
    }
    if (tu.tile[0] * tu.tile[1] * tu.tile[2] != kWG * tu.ppl || tu.tile[0] < 1 || tu.tile[1] < 1 ||
        tu.tile[2] < 1) {
        set_error("tuning: tile extents must multiply to 256 * points-per-lane");
        return ST_ARG;
    }
    for (int c = 0; c < 3; c++) {
        const int *ct = tu.class_tile[c];
        if (ct[0] == 0 && ct[1] == 0 && ct[2] == 0) continue;
        if (ct[0] < 1 || ct[1] < 1 || ct[2] < 1 || ct[0] * ct[1] * ct[2] != kWG * tu.ppl) {
            set_error("tuning: a class tile must be all zero or multiply to 256 * points-per-lane");
            return ST_ARG;
        }
    }
    if (!(tu.class_spacing[0] > 0.0) || !(tu.class_spacing[1] >= tu.class_spacing[0]) ||
        !(tu.class_spacing[1] < 1e6)) {
        set_error("tuning: class spacings must satisfy 0 < h0 <= h1");
        return ST_ARG;
    }
    if (tu.cap < 2 * kWG || tu.cap > 4800) {
        set_error("tuning: LDS list capacity must be in [512, 4800] records");
        return ST_ARG;
    }
    return ST_OK;
}

int widom_prepare(Context *ctx, FrameworkDev *fw, uint64_t n_ins, WidomBuffers *wb,
                  std::vector<void *> *owned) {
    if (n_ins == 0 || n_ins >= (1ULL << 32)) {
        set_error("Widom insertion count must be in [1, 2^32)");
        return ST_ARG;
    }
    choose_widom_bins(fw, n_ins, ctx->tuning, wb->nbins);
    const long long nb = (long long)wb->nbins[0] * wb->nbins[1] * wb->nbins[2];
    ST_TRY(pool_alloc(ctx, nb + 1, &wb->count, owned));
    ST_TRY(pool_alloc(ctx, nb + 1, &wb->start, owned));
    ST_TRY(pool_alloc(ctx, nb + 1, &wb->cursor, owned));
    ST_TRY(pool_alloc(ctx, n_ins, &wb->sorted, owned));
    ST_TRY(pool_alloc(ctx, n_ins, &wb->e, owned));
    ST_TRY(pool_alloc(ctx, 1, &wb->d_job, owned));
    return ST_OK;
}

BinJob make_bin_job(const WidomBuffers &wb, uint64_t n_ins, uint64_t seed, uint32_t stream_id) {
    BinJob j;
    std::memset(&j, 0, sizeof(j));
    j.seed = seed;
    j.n = n_ins;
    j.stream = stream_id;
    for (int d = 0; d < 3; d++) j.nb[d] = wb.nbins[d];
    j.count = wb.count;
    j.start = wb.start;
    j.cursor = wb.cursor;
    j.sorted = wb.sorted;
    j.block0 = 0;
    return j;
}

int widom_enqueue_binning(Context *ctx, WidomBuffers &wb, uint64_t n_ins, uint64_t seed,
                          uint32_t stream_id, hipStream_t s) {
    (void)ctx;
    const long long nb = (long long)wb.nbins[0] * wb.nbins[1] * wb.nbins[2];
    wb.job = make_bin_job(wb, n_ins, seed, stream_id);
    HIP_TRY(hipMemcpyAsync(wb.d_job, &wb.job, sizeof(BinJob), hipMemcpyHostToDevice, s));
    HIP_TRY(launch_bin_all(wb.d_job, 1, bin_blocks(n_ins, kBinChunk), kBinChunk, nb, s));
    return ST_OK;
}


size_t framework_near_capacity(int n_atoms, const double t[9], const int rep[3], double rc2) {
    double inv[9];
    invert3(t, inv);
    const double rc = std::sqrt(rc2);
    size_t per = 1;
    for (int d = 0; d < 3; d++) {
        const double ext = rc / perp_width(inv, d) + 1e-7;
        const long long k =
            std::min<long long>(2LL * rep[d] + 1, (long long)std::floor(1.0 + 2.0 * ext) + 1);
        per *= (size_t)std::max<long long>(k, 1);
    }
    return std::max<size_t>(1, (size_t)std::max(n_atoms, 0) * per);
}

// ---------------------------------------------------------------------------
// Plain path
// ---------------------------------------------------------------------------

size_t GridDev::bytes() const {
    return (size_t)n[0] * n[1] * n[2] * (precision == 1 ? sizeof(double) : sizeof(float));
}

GridDev::~GridDev() {
    if (ctx) ctx->pool.free(d_data);
}

int grid_compute(Context *ctx, FrameworkDev *fw, const int n[3], int precision, void *host_out,
                 GridDev **resident) {
    if (!ctx || !fw || !n) {
        set_error("grid_compute: null argument");
        return ST_ARG;
    }
    ST_TRY(check_grid_dims(n));
    ST_TRY(check_tuning(ctx->tuning));
    if (precision != 0 && precision != 1) {
        set_error("grid_compute: precision must be 0 (FP32) or 1 (FP64)");
        return ST_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    GridDev *g = new GridDev();
    g->ctx = ctx;
    g->fw = fw;
    std::memcpy(g->n, n, sizeof(g->n));
    g->precision = precision;
    g->d_data = ctx->pool.alloc(g->bytes());
    if (!g->d_data) {
        delete g;
        return ST_NOMEM;
    }
    Scoped tmp{ctx, {}};
    const int math = precision == 1 ? MATH_FP64 : MATH_MIXED;
    JobDev job = make_grid_job(fw, n, precision, g->d_data, ctx->tuning, math);
    JobDev *d_job = nullptr;
    int st = pool_alloc(ctx, 1, &d_job, &tmp.ptrs);
    if (st) {
        delete g;
        return st;
    }
    hipStream_t s = ctx->stream;
    auto fail = [&](int code) {
        delete g;
        return code;
    };
    if (hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(ST_HIP);
    if ((st = run_jobs(ctx, math, d_job, 1, job.n_items, ctx->d_counter))) return fail(st);
    if (host_out &&
        hipMemcpyAsync(host_out, g->d_data, g->bytes(), hipMemcpyDeviceToHost, s) != hipSuccess) {
        set_error("grid_compute: device-to-host copy failed");
        return fail(ST_HIP);
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_error(std::string("grid_compute: ") + hipGetErrorString(hipGetLastError()));
        return fail(ST_HIP);
    }
    if (resident) *resident = g;
    else delete g;
    return ST_OK;
}

int widom_run(Context *ctx, FrameworkDev *fw, uint64_t n_ins, double temperature, uint64_t seed,
              WidomResult *out, double *energies_host) {
    if (!ctx || !fw || !out) {
        set_error("widom: null argument");
        return ST_ARG;
    }
    if (!(temperature > 0.0) || !std::isfinite(temperature)) {
        set_error("widom: temperature must be positive");
        return ST_ARG;
    }
    ST_TRY(check_tuning(ctx->tuning));
    HIP_TRY(hipSetDevice(ctx->device));
    Scoped tmp{ctx, {}};
    WidomBuffers wb;
    ST_TRY(widom_prepare(ctx, fw, n_ins, &wb, &tmp.ptrs));
    // energies returned to the caller are computed in full; otherwise only the
    // sums are, and saturated insertions may stop early
    JobDev job = make_widom_job(fw, n_ins, seed, wb, ctx->tuning, kMinGridItems,
                                energies_host ? 0.0 : temperature);
    JobDev *d_job = nullptr;
    ST_TRY(pool_alloc(ctx, 1, &d_job, &tmp.ptrs));
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s));
    ST_TRY(widom_enqueue_binning(ctx, wb, n_ins, seed, 0u, s));
    ST_TRY(run_jobs(ctx, MATH_MIXED, d_job, 1, job.n_items, ctx->d_counter));
    HIP_TRY(launch_widom_partials(wb.e, (long long)n_ins, temperature, ctx->d_partials, s));
    HIP_TRY(launch_widom_final(ctx->d_partials, ctx->d_sums, s));
    double sums[2];
    HIP_TRY(hipMemcpyAsync(sums, ctx->d_sums, sizeof(sums), hipMemcpyDeviceToHost, s));
    if (energies_host)
        HIP_TRY(hipMemcpyAsync(energies_host, wb.e, sizeof(double) * n_ins, hipMemcpyDeviceToHost,
                               s));
    HIP_TRY(hipStreamSynchronize(s));
    out->sum_b = sums[0];
    out->sum_eb = sums[1];
    out->n = n_ins;
    return ST_OK;
}

int count_incut_grid(Context *ctx, FrameworkDev *fw, const int n[3], unsigned long long *out,
                     uint32_t *per_point_host) {
    ST_TRY(check_grid_dims(n));
    ST_TRY(check_tuning(ctx->tuning));
    HIP_TRY(hipSetDevice(ctx->device));
    Scoped tmp{ctx, {}};
    uint32_t *d_cnt = nullptr;
    unsigned long long *d_sum = nullptr;
    JobDev *d_job = nullptr;
    const long long p = (long long)n[0] * n[1] * n[2];
    ST_TRY(pool_alloc(ctx, p, &d_cnt, &tmp.ptrs));
    ST_TRY(pool_alloc(ctx, 1, &d_sum, &tmp.ptrs));
    ST_TRY(pool_alloc(ctx, 1, &d_job, &tmp.ptrs));
    JobDev job = make_grid_job(fw, n, 0, d_cnt, ctx->tuning, MATH_COUNT);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s));
    ST_TRY(run_jobs(ctx, MATH_COUNT, d_job, 1, job.n_items, ctx->d_counter));
    HIP_TRY(launch_fill_u64(d_sum, 0ull, 1, s));
    HIP_TRY(launch_sum_u32(d_cnt, p, d_sum, s));
    HIP_TRY(hipMemcpyAsync(out, d_sum, sizeof(*out), hipMemcpyDeviceToHost, s));
    if (per_point_host)
        HIP_TRY(hipMemcpyAsync(per_point_host, d_cnt, sizeof(uint32_t) * p,
                               hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ST_OK;
}

int accessibility_grid(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                       const int n[3], double *host_out) {
    if (!ctx || (n_atoms > 0 && !atoms) || n_atoms < 0 || !t || !n || !host_out) {
        set_error("accessibility_grid: bad arguments");
        return ST_ARG;
    }
    ST_TRY(check_grid_dims(n));
    ST_TRY(check_tuning(ctx->tuning));
    // candidates: images [-1, 1]^3 within sigma_max of the cell
    // (src/computegridsheet.h:131-133 loops those images whatever the
    // replication factors; an atom farther than sigma_max overlaps nothing)
    double sig2_max = 0.0;
    for (int i = 0; i < n_atoms; i++) sig2_max = std::max(sig2_max, atoms[i].sig2);
    const int rep1[3] = {1, 1, 1};
    FrameworkDev *fw = nullptr;
    ST_TRY(framework_upload(ctx, atoms, n_atoms, t, rep1, std::max(sig2_max, 1e-300), &fw));
    std::unique_ptr<FrameworkDev> owner(fw);
    HIP_TRY(hipSetDevice(ctx->device));
    Scoped tmp{ctx, {}};
    const long long p = (long long)n[0] * n[1] * n[2];
    double *d_out = nullptr;
    JobDev *d_job = nullptr;
    ST_TRY(pool_alloc(ctx, p, &d_out, &tmp.ptrs));
    ST_TRY(pool_alloc(ctx, 1, &d_job, &tmp.ptrs));
    JobDev job = make_grid_job(fw, n, 1, d_out, ctx->tuning, MATH_OVERLAP);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s));
    ST_TRY(run_jobs(ctx, MATH_OVERLAP, d_job, 1, job.n_items, ctx->d_counter));
    HIP_TRY(hipMemcpyAsync(host_out, d_out, sizeof(double) * p, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ST_OK;
}

int count_incut_widom(Context *ctx, FrameworkDev *fw, uint64_t n_ins, uint64_t seed,
                      unsigned long long *out) {
    ST_TRY(check_tuning(ctx->tuning));
    HIP_TRY(hipSetDevice(ctx->device));
    Scoped tmp{ctx, {}};
    WidomBuffers wb;
    ST_TRY(widom_prepare(ctx, fw, n_ins, &wb, &tmp.ptrs));
    unsigned long long *d_sum = nullptr;
    JobDev *d_job = nullptr;
    ST_TRY(pool_alloc(ctx, 1, &d_sum, &tmp.ptrs));
    ST_TRY(pool_alloc(ctx, 1, &d_job, &tmp.ptrs));
    JobDev job = make_widom_job(fw, n_ins, seed, wb, ctx->tuning, kMinGridItems);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s));
    ST_TRY(widom_enqueue_binning(ctx, wb, n_ins, seed, 0u, s));
    ST_TRY(run_jobs(ctx, MATH_COUNT, d_job, 1, job.n_items, ctx->d_counter));
    HIP_TRY(launch_fill_u64(d_sum, 0ull, 1, s));
    HIP_TRY(launch_sum_f64_counts(wb.e, (long long)n_ins, d_sum, s));
    HIP_TRY(hipMemcpyAsync(out, d_sum, sizeof(*out), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ST_OK;
}

// ---------------------------------------------------------------------------
// Batched path
// ---------------------------------------------------------------------------

Batch::~Batch() {
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (GridDev *g : grids) delete g;
    if (h_sums) (void)hipHostFree(h_sums);
    if (ctx)
        for (void *p : owned) ctx->pool.free(p);
}

namespace {

// Expected in-cutoff pairs per point: the ordering key of the LPT schedule.
double job_cost_per_item(const FrameworkDev *fw, const JobDev &j, const Tuning &tu) {
    const double pi = 3.14159265358979323846;
    const double dens = fw->n_atoms / std::max(fw->volume, 1e-9);
    double pts;
    if (j.kind == JOB_GRID)
        pts = (double)j.tile[0] * j.tile[1] * j.tile[2] * j.region[0] * j.region[1] * j.region[2];
    else pts = (double)tu.widom_per_bin * j.region[0] * j.region[1] * j.region[2];
    return pts * dens * (4.0 / 3.0) * pi * fw->rc2 * fw->rc;
}

}  // namespace

int Batch::record(hipStream_t s) {
    // every Widom task's binning in one launch per phase
    HIP_TRY(launch_bin_all(d_bin_jobs, (int)bin_jobs.size(), bin_blocks_total, kBinChunk,
                           std::max<long long>(bin_max, 1), s));
    // jobs are partitioned: [0, n_mixed) mixed math, [n_mixed, njobs) FP64 grids
    int n_mixed = 0;
    long long items_mixed = 0, items_fp64 = 0;
    for (const JobDev &j : jobs) {
        if (j.kind == JOB_GRID && j.mode == OUT_F64) items_fp64 += j.n_items;
        else {
            n_mixed++;
            items_mixed += j.n_items;
        }
    }
    const bool ev_on = timing && !use_graph;
    timed_launches = 0;
    HIP_TRY(launch_fill_u64(d_counter, 0ull, 2, s));
    if (items_mixed > 0) {
        if (ev_on) HIP_TRY(hipEventRecord(ev[0], s));
        HIP_TRY(launch_tiles(MATH_MIXED, ctx->tuning.ppl, d_jobs, n_mixed, items_mixed, d_counter,
                             persistent_wgs(ctx, MATH_MIXED, items_mixed), ctx->tuning.cap,
                             ctx->d_scratch, kRegionRecords, s));
        if (ev_on) HIP_TRY(hipEventRecord(ev[1], s));
        timed_launches |= 1;
    }
    if (items_fp64 > 0) {
        if (ev_on) HIP_TRY(hipEventRecord(ev[2], s));
        HIP_TRY(launch_tiles(MATH_FP64, ctx->tuning.ppl, d_jobs + n_mixed,
                             (int)jobs.size() - n_mixed, items_fp64, d_counter + 1,
                             persistent_wgs(ctx, MATH_FP64, items_fp64), ctx->tuning.cap,
                             ctx->d_scratch, kRegionRecords, s));
        if (ev_on) HIP_TRY(hipEventRecord(ev[3], s));
        timed_launches |= 2;
    }
    // every Widom run's sums in two launches (same per-run order as one run alone)
    HIP_TRY(launch_widom_reduce_many(d_reduce, (int)reduce.size(), s));
    return ST_OK;
}

int batch_create(Context *ctx, const Task *tasks, int n, bool use_graph, Batch **out) {
    if (!ctx || (!tasks && n > 0) || n < 0 || !out) {
        set_error("batch_create: bad arguments");
        return ST_ARG;
    }
    ST_TRY(check_tuning(ctx->tuning));
    HIP_TRY(hipSetDevice(ctx->device));
    Batch *b = new Batch();
    b->ctx = ctx;
    b->use_graph = use_graph;
    b->tasks.assign(tasks, tasks + n);
    b->grids.assign(n, nullptr);
    b->wbufs.assign(n, WidomBuffers());
    auto fail = [&](int st) {
        delete b;
        return st;
    };
    std::vector<std::pair<double, JobDev>> mixed, fp64;
    for (int t = 0; t < n; t++) {
        const Task &tk = tasks[t];
        if (!tk.fw) {
            set_error("batch_create: task without framework");
            return fail(ST_ARG);
        }
        const bool has_grid = tk.grid_n[0] || tk.grid_n[1] || tk.grid_n[2];
        if (has_grid) {
            int st = check_grid_dims(tk.grid_n);
            if (st) return fail(st);
            if (tk.precision != 0 && tk.precision != 1) {
                set_error("batch_create: precision must be 0 or 1");
                return fail(ST_ARG);
            }
            GridDev *g = new GridDev();
            g->ctx = ctx;
            g->fw = tk.fw;
            std::memcpy(g->n, tk.grid_n, sizeof(g->n));
            g->precision = tk.precision;
            g->d_data = ctx->pool.alloc(g->bytes());
            if (!g->d_data) {
                delete g;
                return fail(ST_NOMEM);
            }
            b->grids[t] = g;
            const int math = tk.precision == 1 ? MATH_FP64 : MATH_MIXED;
            JobDev j = make_grid_job(tk.fw, tk.grid_n, tk.precision, g->d_data, ctx->tuning, math);
            (math == MATH_FP64 ? fp64 : mixed)
                .push_back({job_cost_per_item(tk.fw, j, ctx->tuning), j});
        }
        if (tk.n_ins > 0) {
            if (!(tk.temperature > 0.0)) {
                set_error("batch_create: temperature must be positive");
                return fail(ST_ARG);
            }
            int st = widom_prepare(ctx, tk.fw, tk.n_ins, &b->wbufs[t], &b->owned);
            if (st) return fail(st);
            BinJob bj = make_bin_job(b->wbufs[t], tk.n_ins, tk.seed, 0u);
            bj.block0 = b->bin_blocks_total;
            b->bin_blocks_total += bin_blocks(tk.n_ins, kBinChunk);
            b->bin_max = std::max<long long>(
                b->bin_max, (long long)bj.nb[0] * bj.nb[1] * bj.nb[2]);
            b->bin_jobs.push_back(bj);
            // many jobs share the launch: a Widom job needs only its share of
            // the device's work items
            JobDev j = make_widom_job(tk.fw, tk.n_ins, tk.seed, b->wbufs[t], ctx->tuning,
                                      std::max<long long>(1, kMinGridItems / std::max(n, 1)),
                                      tk.temperature);
            mixed.push_back({job_cost_per_item(tk.fw, j, ctx->tuning), j});
        }
    }
    // Longest-processing-time order: the most expensive items are handed out first.
    auto by_cost = [](const std::pair<double, JobDev> &a, const std::pair<double, JobDev> &c) {
        return a.first > c.first;
    };
    std::stable_sort(mixed.begin(), mixed.end(), by_cost);
    std::stable_sort(fp64.begin(), fp64.end(), by_cost);
    long long begin = 0;
    for (auto &pj : mixed) {
        pj.second.item_begin = begin;
        begin += pj.second.n_items;
        b->jobs.push_back(pj.second);
    }
    begin = 0;
    for (auto &pj : fp64) {
        pj.second.item_begin = begin;
        begin += pj.second.n_items;
        b->jobs.push_back(pj.second);
    }
    int st = ST_OK;
    if ((st = pool_alloc(ctx, std::max<size_t>(b->jobs.size(), 1), &b->d_jobs, &b->owned)) ||
        (st = pool_alloc(ctx, (size_t)std::max(n, 1) * kReduceBlocks, &b->d_partials, &b->owned)) ||
        (st = pool_alloc(ctx, (size_t)std::max(n, 1) * 2, &b->d_sums, &b->owned)) ||
        (st = pool_alloc(ctx, 2, &b->d_counter, &b->owned)) ||
        (st = pool_alloc(ctx, std::max<size_t>(b->bin_jobs.size(), 1), &b->d_bin_jobs,
                         &b->owned)) ||
        (st = pool_alloc(ctx, (size_t)std::max(n, 1), &b->d_reduce, &b->owned)))
        return fail(st);
    for (int t = 0; t < n; t++)
        if (b->tasks[t].n_ins > 0)
            b->reduce.push_back({b->wbufs[t].e, (long long)b->tasks[t].n_ins,
                                 b->tasks[t].temperature, b->d_partials + (size_t)t * kReduceBlocks,
                                 b->d_sums + 2 * (size_t)t});
    if (hipHostMalloc((void **)&b->h_sums, sizeof(double) * 2 * std::max(n, 1)) != hipSuccess) {
        set_error("batch_create: pinned host allocation failed");
        return fail(ST_NOMEM);
    }
    if ((!b->jobs.empty() &&
         hipMemcpyAsync(b->d_jobs, b->jobs.data(), sizeof(JobDev) * b->jobs.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (!b->bin_jobs.empty() &&
         hipMemcpyAsync(b->d_bin_jobs, b->bin_jobs.data(), sizeof(BinJob) * b->bin_jobs.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (!b->reduce.empty() &&
         hipMemcpyAsync(b->d_reduce, b->reduce.data(), sizeof(WidomReduce) * b->reduce.size(),
                        hipMemcpyHostToDevice, ctx->stream) != hipSuccess)) {
        set_error("batch_create: job upload failed");
        return fail(ST_HIP);
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("batch_create: stream synchronize failed");
        return fail(ST_HIP);
    }
    *out = b;
    return ST_OK;
}

int batch_run(Batch *b) {
    if (!b) {
        set_error("batch_run: null batch");
        return ST_ARG;
    }
    Context *ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    if (!b->use_graph) return b->record(ctx->stream);
    if (!b->exec) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int st = b->record(ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st != ST_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        HIP_TRY(e);
        b->graph = g;
        HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipGraphLaunch(b->exec, ctx->stream));
    return ST_OK;
}

int batch_wait(Batch *b) {
    if (!b) {
        set_error("batch_wait: null batch");
        return ST_ARG;
    }
    Context *ctx = b->ctx;
    const size_t n = b->tasks.size();
    if (n > 0)
        HIP_TRY(hipMemcpyAsync(b->h_sums, b->d_sums, sizeof(double) * 2 * n,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ST_OK;
}

int batch_set_timing(Batch *b, bool on) {
    if (!b) {
        set_error("batch_set_timing: null batch");
        return ST_ARG;
    }
    if (on && !b->ev[0])
        for (hipEvent_t &e : b->ev) HIP_TRY(hipEventCreate(&e));
    b->timing = on;
    return ST_OK;
}

int batch_kernel_ms(Batch *b, double *ms, int *launches) {
    if (!b || !ms) {
        set_error("batch_kernel_ms: null argument");
        return ST_ARG;
    }
    *ms = 0.0;
    int n = 0;
    if (b->timing && !b->use_graph) {
        for (int k = 0; k < 2; k++)
            if (b->timed_launches & (1 << k)) {
                float t = 0.0f;
                HIP_TRY(hipEventElapsedTime(&t, b->ev[2 * k], b->ev[2 * k + 1]));
                *ms += t;
                n++;
            }
    }
    if (launches) *launches = n;
    return ST_OK;
}

int batch_result(Batch *b, int i, WidomResult *out, GridDev **grid) {
    if (!b || i < 0 || i >= (int)b->tasks.size()) {
        set_error("batch_result: index out of range");
        return ST_ARG;
    }
    if (out) {
        out->sum_b = b->h_sums[2 * i];
        out->sum_eb = b->h_sums[2 * i + 1];
        out->n = b->tasks[i].n_ins;
    }
    if (grid) *grid = b->grids[i];
    return ST_OK;
}

}  // namespace er
