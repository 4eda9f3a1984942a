// er_pockets_host.cpp -- pocket blocking and grid-interpolated Widom on
// resident FP32 grids (src/pocketblocking.cc:126-391, src/gcmc.cc:145-177).
#include <array>
#include <cmath>
#include <cstring>
#include <sstream>
#include <vector>

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

int pockets_label(Context *ctx, GridDev *g, double temperature, bool block, int *n_pockets,
                  int32_t *seg_host) {
    if (!ctx || !g || !n_pockets) {
        set_error("pockets: null argument");
        return ST_ARG;
    }
    if (g->precision != 0) {
        set_error("pockets: the grid must be FP32");
        return ST_ARG;
    }
    if (!(temperature > 0.0)) {
        set_error("pockets: temperature must be positive");
        return ST_ARG;
    }
    const long long nv = (long long)g->n[0] * g->n[1] * g->n[2];
    if (nv >= (1LL << 31) - 1) {
        set_error("pockets: grid too large for 32-bit voxel indices");
        return ST_ARG;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    Scratch sc{ctx, {}};
    const long long nfaces = (long long)g->n[1] * g->n[2] + (long long)g->n[0] * g->n[2] +
                             (long long)g->n[0] * g->n[1];
    const long long nblk = (nv + kScanChunk - 1) / kScanChunk;
    int *parent = sc.get<int>(nv), *seg = sc.get<int>(nv), *links = sc.get<int>(3 * nfaces);
    uint32_t *is_root = sc.get<uint32_t>(nv), *rank = sc.get<uint32_t>(nv + 1);
    uint32_t *bsum = sc.get<uint32_t>(nblk + 2), *boff = sc.get<uint32_t>(nblk + 2);
    if (!parent || !seg || !links || !is_root || !rank || !bsum || !boff) return ST_NOMEM;
    hipStream_t s = ctx->stream;
    const float *e = (const float *)g->d_data;
    // src/pocketblocking.cc:130: E >= 15 T is inaccessible
    HIP_TRY(launch_cc_init(e, nv, 15.0 * temperature, parent, s));
    HIP_TRY(launch_cc_hook(g->n, parent, s));
    HIP_TRY(launch_cc_flatten(nv, parent, is_root, s));
    HIP_TRY(launch_scan_large(is_root, nv, rank, bsum, boff, s));
    HIP_TRY(launch_cc_segments(nv, parent, rank, seg, s));
    HIP_TRY(launch_cc_faces(g->n, seg, links, s));
    std::vector<int> hl(3 * nfaces);
    uint32_t nseg = 0;
    HIP_TRY(hipMemcpyAsync(hl.data(), links, sizeof(int) * 3 * nfaces, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&nseg, rank + nv, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<uint8_t> pocket;
    const int np = classify_pockets((int)nseg, hl.data(), nfaces, &pocket);
    if (np < 0) {
        set_error("pockets: inconsistent face links");
        return ST_STATE;
    }
    if (block && np > 0) {
        uint8_t *d_pocket = sc.get<uint8_t>(pocket.size());
        if (!d_pocket) return ST_NOMEM;
        HIP_TRY(hipMemcpyAsync(d_pocket, pocket.data(), pocket.size(), hipMemcpyHostToDevice, s));
        HIP_TRY(launch_cc_block(nv, seg, d_pocket, (float *)g->d_data, s));
    }
    if (seg_host)
        HIP_TRY(hipMemcpyAsync(seg_host, seg, sizeof(int) * nv, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *n_pockets = np;
    return ST_OK;
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
    double *e = sc.get<double>(n_ins);
    if (!e) return ST_NOMEM;
    hipStream_t s = ctx->stream;
    HIP_TRY(launch_widom_grid((const float *)g->d_data, g->n, seed, 1u, n_ins, e, s));
    HIP_TRY(launch_widom_partials(e, (long long)n_ins, temperature, ctx->d_partials, s));
    HIP_TRY(launch_widom_final(ctx->d_partials, ctx->d_sums, s));
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

}  // namespace er
