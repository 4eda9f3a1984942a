// er_pockets_host.cpp -- pocket blocking and grid-interpolated Widom on
// resident FP32 grids (src/pocketblocking.cc:126-391, src/gcmc.cc:145-177).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <sstream>
#include <thread>
#include <vector>

#include "er_kernels.h"
#include "er_pockets.h"
#include "er_runtime.h"

namespace er {

namespace {

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::ostringstream os_;                                                        \
            os_ << "HIP error " << hipGetErrorName(e_) << " at " << __FILE__ << ":"        \
                << __LINE__ << " in " << #expr;                                            \
            set_error(os_.str());                                                          \
            return ST_HIP;                                                                 \
        }                                                                                  \
    } while (0)

#define ST_TRY_P(expr)              \
    do {                            \
        int st_ = (expr);           \
        if (st_ != ST_OK) return st_; \
    } while (0)

struct Scratch {
    Context *ctx;
    std::vector<void *> ptrs;
    template <class T>
    T *get(size_t count) {
        void *p = ctx->pool.alloc(count * sizeof(T) > 0 ? count * sizeof(T) : 16);
        if (p) ptrs.push_back(p);
        return (T *)p;
    }
    ~Scratch() {
        for (void *p : ptrs) ctx->pool.free(p);
    }
};

// Offset-carrying union-find over segments.  A link (from a to b, axis d)
// says: walking out of a through the +d face of the cell enters b in the
// next cell, uc(b) = uc(a) + e_d.  Two paths between the same segments with
// different net translations (or a link of a segment to itself) mean the
// connected part of the graph wraps around the crystal: a channel.  This is
// the property src/pocketblocking.cc:39-100 decides by depth-first search.
struct OffsetUnionFind {
    std::vector<int> parent, size;
    std::vector<std::array<int, 3>> off;  // position relative to parent
    std::vector<uint8_t> channel;
    explicit OffsetUnionFind(int n) : parent(n), size(n, 1), off(n, {0, 0, 0}), channel(n, 0) {
        for (int i = 0; i < n; i++) parent[i] = i;
    }
    int find(int x, std::array<int, 3> &pos) {
        pos = {0, 0, 0};
        int r = x;
        while (parent[r] != r) {
            for (int d = 0; d < 3; d++) pos[d] += off[r][d];
            r = parent[r];
        }
        // compress: point every node on the path straight at the root
        std::array<int, 3> rem = pos;
        int y = x;
        while (parent[y] != y) {
            const int next = parent[y];
            const std::array<int, 3> oy = off[y];
            off[y] = rem;
            parent[y] = r;
            for (int d = 0; d < 3; d++) rem[d] -= oy[d];
            y = next;
        }
        return r;
    }
    void link(int a, int b, int axis) {
        std::array<int, 3> pa, pb;
        const int ra = find(a, pa), rb = find(b, pb);
        std::array<int, 3> delta = {0, 0, 0};
        delta[axis] = 1;
        if (ra == rb) {
            for (int d = 0; d < 3; d++)
                if (pb[d] - pa[d] != delta[d]) channel[ra] = 1;
            return;
        }
        // pos(b) must equal pos(a) + delta
        if (size[ra] >= size[rb]) {
            parent[rb] = ra;
            for (int d = 0; d < 3; d++) off[rb][d] = pa[d] + delta[d] - pb[d];
            size[ra] += size[rb];
            channel[ra] |= channel[rb];
        } else {
            parent[ra] = rb;
            for (int d = 0; d < 3; d++) off[ra][d] = pb[d] - delta[d] - pa[d];
            size[rb] += size[ra];
            channel[rb] |= channel[ra];
        }
    }
};

}  // namespace

int classify_pockets(int nseg, const int *links, long long nlinks, std::vector<uint8_t> *pocket) {
    OffsetUnionFind uf(nseg > 0 ? nseg : 1);
    for (long long q = 0; q < nlinks; q++) {
        const int a = links[3 * q], b = links[3 * q + 1], dir = links[3 * q + 2];
        if (a < 0) continue;
        if (a >= nseg || b < 0 || b >= nseg || dir < 1 || dir > 3) return -1;
        uf.link(a, b, dir - 1);
    }
    pocket->assign(nseg > 0 ? nseg : 1, 0);
    int np = 0;
    for (int s = 0; s < nseg; s++) {
        std::array<int, 3> pos;
        const int r = uf.find(s, pos);
        (*pocket)[s] = uf.channel[r] ? 0 : 1;
        np += (*pocket)[s];
    }
    return np;
}

int grid_upload(Context *ctx, const float *host, const int n[3], GridDev **out) {
    if (!ctx || !host || !n || !out) {
        set_error("grid_upload: null argument");
        return ST_ARG;
    }
    for (int d = 0; d < 3; d++)
        if (n[d] < 2 || n[d] > 65536) {
            set_error("grid dimensions must be in [2, 65536]");
            return ST_ARG;
        }
    HIP_TRY(hipSetDevice(ctx->device));
    GridDev *g = new GridDev();
    g->ctx = ctx;
    std::memcpy(g->n, n, sizeof(g->n));
    g->precision = 0;
    g->d_data = ctx->pool.alloc(g->bytes());
    if (!g->d_data) {
        delete g;
        return ST_NOMEM;
    }
    if (hipMemcpyAsync(g->d_data, host, g->bytes(), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        delete g;
        set_error("grid_upload: copy failed");
        return ST_HIP;
    }
    *out = g;
    return ST_OK;
}

namespace {

// Voxels labelled per launch sequence (about 4.3 B of scratch per voxel).
constexpr long long kCCChunkVoxels = 1LL << 28;

int label_chunk(Context *ctx, GridDev *const *grids, int ng, double temperature, bool block,
                int *n_pockets, int32_t *const *seg_host) {
    hipStream_t s = ctx->stream;
    Scratch sc{ctx, {}};
    std::vector<CCGrid> tab(ng);
    std::vector<long long> faces(ng);
    long long nv_sum = 0, nw_sum = 0, face_sum = 0, blocks = 0, bricks = 0, fblocks = 0;
    for (int g = 0; g < ng; g++) {
        const int *n = grids[g]->n;
        CCGrid &c = tab[g];
        std::memset(&c, 0, sizeof(c));
        c.e = (float *)grids[g]->d_data;
        std::memcpy(c.n, n, sizeof(c.n));
        c.nv = (long long)n[0] * n[1] * n[2];
        c.nw = (c.nv + 63) / 64;
        c.blk0 = blocks;
        c.bblk0 = bricks;
        c.fblk0 = fblocks;
        c.seg0 = nv_sum;
        bricks += cc_bricks(n);
        fblocks += cc_face_blocks(n);
        faces[g] = (long long)n[1] * n[2] + (long long)n[0] * n[2] + (long long)n[0] * n[1];
        blocks += cc_blocks(c.nv);
        nv_sum += c.nv;
        nw_sum += c.nw + 1;
        face_sum += faces[g];
    }
    int *parent = sc.get<int>(nv_sum), *links = sc.get<int>(3 * face_sum),
        *packed = sc.get<int>(3 * face_sum);
    unsigned long long *mask = sc.get<unsigned long long>(nw_sum);
    uint32_t *wpre = sc.get<uint32_t>(nw_sum);
    unsigned int *meta = sc.get<unsigned int>(4 * (size_t)ng);
    CCGrid *d_tab = sc.get<CCGrid>(ng);
    if (!parent || !links || !packed || !mask || !wpre || !meta || !d_tab) return ST_NOMEM;
    long long ov = 0, ow = 0, of = 0;
    for (int g = 0; g < ng; g++) {
        tab[g].parent = parent + ov;
        tab[g].mask = mask + ow;
        tab[g].wpre = wpre + ow;
        tab[g].links = links + 3 * of;
        ov += tab[g].nv;
        ow += tab[g].nw + 1;
        of += faces[g];
    }
    HIP_TRY(hipMemcpyAsync(d_tab, tab.data(), sizeof(CCGrid) * ng, hipMemcpyHostToDevice, s));
    // src/pocketblocking.cc:130: E >= 15 T is inaccessible
    HIP_TRY(launch_cc_label(d_tab, ng, blocks, bricks, fblocks, 15.0 * temperature, packed, meta,
                            s));
    // One pinned staging buffer: the per-grid meta and the first kLinkPrefix
    // packed links come back with one synchronisation; a larger link set
    // takes a second copy.
    constexpr long long kLinkPrefix = 1 << 18;
    const long long prefix = std::min<long long>(kLinkPrefix, std::max<long long>(face_sum, 1));
    const size_t meta_bytes = sizeof(unsigned int) * 4 * (size_t)ng;
    const size_t stage_bytes = ((meta_bytes + 15) & ~(size_t)15) + sizeof(int) * 3 * prefix;
    char *stage = (char *)ctx->host_stage(stage_bytes);
    if (!stage) return ST_NOMEM;
    unsigned int *hmeta = (unsigned int *)stage;
    int *hl = (int *)(stage + ((meta_bytes + 15) & ~(size_t)15));
    HIP_TRY(hipMemcpyAsync(hmeta, meta, meta_bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hl, packed, sizeof(int) * 3 * prefix, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    long long nlinks = 0;
    for (int g = 0; g < ng; g++) nlinks += hmeta[4 * g];
    std::vector<int> hl_big;
    if (nlinks > prefix) {
        hl_big.resize(3 * (size_t)nlinks);
        HIP_TRY(hipMemcpyAsync(hl_big.data(), packed, sizeof(int) * 3 * nlinks,
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        hl = hl_big.data();
    }
    // classify every grid (independent; a large link set is spread over host
    // threads, a small one is cheaper than starting them)
    std::vector<std::vector<uint8_t>> pocket(ng);
    std::vector<int> np(ng, 0);
    auto work = [&](int g) {
        np[g] = classify_pockets((int)hmeta[4 * g + 1], hl + 3 * (size_t)hmeta[4 * g + 2],
                                 hmeta[4 * g], &pocket[g]);
    };
    const int nthreads =
        nlinks < 100000
            ? 1
            : std::min<int>(ng, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (nthreads <= 1) {
        for (int g = 0; g < ng; g++) work(g);
    } else {
        std::vector<std::thread> pool;
        std::atomic<int> next{0};
        for (int t = 0; t < nthreads; t++)
            pool.emplace_back([&] {
                for (int g = next++; g < ng; g = next++) work(g);
            });
        for (auto &th : pool) th.join();
    }
    for (int g = 0; g < ng; g++)
        if (np[g] < 0) {
            set_error("pockets: inconsistent face links");
            return ST_STATE;
        }
    bool any_seg = false;
    for (int g = 0; g < ng && seg_host; g++) any_seg |= seg_host[g] != nullptr;
    long long pk_sum = 0;
    for (int g = 0; g < ng; g++)
        if (block && np[g] > 0) pk_sum += (long long)pocket[g].size();
    if (pk_sum > 0) {
        std::vector<uint8_t> hp(pk_sum);
        uint8_t *d_pocket = sc.get<uint8_t>(pk_sum);
        if (!d_pocket) return ST_NOMEM;
        long long o = 0;
        for (int g = 0; g < ng; g++) {
            tab[g].pocket = nullptr;
            if (!(block && np[g] > 0)) continue;
            std::memcpy(hp.data() + o, pocket[g].data(), pocket[g].size());
            tab[g].pocket = d_pocket + o;
            o += (long long)pocket[g].size();
        }
        HIP_TRY(hipMemcpyAsync(d_pocket, hp.data(), pk_sum, hipMemcpyHostToDevice, s));
        // blocking changes energies only; segment ids come from parent / mask
        HIP_TRY(hipMemcpyAsync(d_tab, tab.data(), sizeof(CCGrid) * ng, hipMemcpyHostToDevice, s));
        HIP_TRY(launch_cc_block(d_tab, ng, blocks, s));
    }
    if (any_seg) {
        int *seg = sc.get<int>(nv_sum);
        if (!seg) return ST_NOMEM;
        HIP_TRY(launch_cc_segments(d_tab, ng, blocks, seg, s));
        for (int g = 0; g < ng; g++)
            if (seg_host[g])
                HIP_TRY(hipMemcpyAsync(seg_host[g], seg + tab[g].seg0, sizeof(int) * tab[g].nv,
                                       hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (int g = 0; g < ng; g++) n_pockets[g] = np[g];
    return ST_OK;
}

}  // namespace

int pockets_label_many(Context *ctx, GridDev *const *grids, int ng, double temperature,
                       bool block, int *n_pockets, int32_t *const *seg_host) {
    if (!ctx || !grids || !n_pockets || ng < 0) {
        set_error("pockets: null argument");
        return ST_ARG;
    }
    if (!(temperature > 0.0)) {
        set_error("pockets: temperature must be positive");
        return ST_ARG;
    }
    for (int g = 0; g < ng; g++) {
        if (!grids[g]) {
            set_error("pockets: null grid");
            return ST_ARG;
        }
        if (grids[g]->precision != 0) {
            set_error("pockets: the grid must be FP32");
            return ST_ARG;
        }
        const long long nv = (long long)grids[g]->n[0] * grids[g]->n[1] * grids[g]->n[2];
        if (nv >= (1LL << 31) - 1) {
            set_error("pockets: grid too large for 32-bit voxel indices");
            return ST_ARG;
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    int g0 = 0;
    while (g0 < ng) {
        // grids of this chunk: at least one, then up to the voxel budget
        long long nv = 0;
        int g1 = g0;
        do {
            nv += (long long)grids[g1]->n[0] * grids[g1]->n[1] * grids[g1]->n[2];
            g1++;
        } while (g1 < ng && nv + (long long)grids[g1]->n[0] * grids[g1]->n[1] * grids[g1]->n[2] <=
                                kCCChunkVoxels);
        ST_TRY_P(label_chunk(ctx, grids + g0, g1 - g0, temperature, block, n_pockets + g0,
                             seg_host ? seg_host + g0 : nullptr));
        g0 = g1;
    }
    return ST_OK;
}

int pockets_label(Context *ctx, GridDev *g, double temperature, bool block, int *n_pockets,
                  int32_t *seg_host) {
    if (!g) {
        set_error("pockets: null argument");
        return ST_ARG;
    }
    int32_t *segs[1] = {seg_host};
    return pockets_label_many(ctx, &g, 1, temperature, block, n_pockets, segs);
}

int widom_grid_run(Context *ctx, GridDev *g, uint64_t n_ins, double temperature, uint64_t seed,
                   WidomResult *out, double *energies_host) {
    if (!ctx || !g || !out) {
        set_error("widom_grid: null argument");
        return ST_ARG;
    }
    if (g->precision != 0) {
        set_error("widom_grid: the grid must be FP32");
        return ST_ARG;
    }
    if (n_ins == 0 || n_ins >= (1ULL << 40) || !(temperature > 0.0)) {
        set_error("widom_grid: insertion count in [1, 2^40) and positive temperature required");
        return ST_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    Scratch sc{ctx, {}};
    double *e = energies_host ? sc.get<double>(n_ins) : nullptr;
    double2 *partials = sc.get<double2>(kTriBlocks);
    if ((energies_host && !e) || !partials) return ST_NOMEM;
    hipStream_t s = ctx->stream;
    HIP_TRY(launch_widom_grid((const float *)g->d_data, g->n, seed, 1u, n_ins, temperature, e,
                              partials, s));
    HIP_TRY(launch_widom_final_n(partials, kTriBlocks, ctx->d_sums, s));
    double sums[2];
    HIP_TRY(hipMemcpyAsync(sums, ctx->d_sums, sizeof(sums), hipMemcpyDeviceToHost, s));
    if (energies_host)
        HIP_TRY(hipMemcpyAsync(energies_host, e, sizeof(double) * n_ins, hipMemcpyDeviceToHost,
                               s));
    HIP_TRY(hipStreamSynchronize(s));
    out->sum_b = sums[0];
    out->sum_eb = sums[1];
    out->n = n_ins;
    return ST_OK;
}

int grid_energies(Context *ctx, GridDev *g, const double *frac_host, uint64_t n,
                  double *out_host) {
    if (!ctx || !g || (n && (!frac_host || !out_host))) {
        set_error("grid_energies: null argument");
        return ST_ARG;
    }
    if (n == 0) return ST_OK;
    if (n >= (1ULL << 40)) {
        set_error("grid_energies: position count must be below 2^40");
        return ST_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    Scratch sc{ctx, {}};
    double *f = sc.get<double>(3 * n), *e = sc.get<double>(n);
    if (!f || !e) return ST_NOMEM;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(f, frac_host, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_grid_energies(g->d_data, g->precision, g->n, f, (long long)n, e, s));
    HIP_TRY(hipMemcpyAsync(out_host, e, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ST_OK;
}

}  // namespace er
