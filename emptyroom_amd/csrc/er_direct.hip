// er_direct.hip -- the reference's direct method on the GPU: every (point,
// image, atom) triple in the order of src/computegridsheet.h:54-87 and
// src/doHenry.h:69-99, with the reference's arithmetic in FP64 (image
// position from fractional coordinates, displacement, r^2, clamp, strict
// cutoff, sigma^2 / r^2 by IEEE division).  No candidate lists, no pruning,
// no mixed precision: this is the baseline the tile kernel is measured
// against, and a second GPU implementation that must equal the oracle bit
// for bit.
//
// One point (or insertion) per lane; a 256-lane block builds 256 image atoms
// at a time in LDS (one per lane, from the fractional coordinates and the
// image indices), then every lane walks them in order.  Implicit FMA
// contraction is off, so every operation rounds as written.
#pragma clang fp contract(off)

#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "er_kernels.h"
#include "er_runtime.h"

namespace er {

namespace {

constexpr int kDirectWG = 256;

struct DirectJob {
    double t[9];
    int rep[3];
    int n_atoms;
    long long n_q;  // images x atoms
    double rc2, min_r2;
    const Particle *atoms;  // original order
};

// src/computegridsheet.h:21-28
__device__ __forceinline__ void frac_to_cart(const double (&t)[9], double xf, double yf,
                                             double zf, double &x, double &y, double &z) {
    x = t[0] * xf + t[1] * yf + t[2] * zf;
    y = t[3] * xf + t[4] * yf + t[5] * zf;
    z = t[6] * xf + t[7] * yf + t[8] * zf;
}

// Energy at Cartesian (x, y, z): images outer (a fastest, then b, c as in
// the reference's nested loops with c innermost), atoms inner.  Every lane of
// the block calls this (the staging uses barriers); q runs over image-major
// (image, atom) pairs.
__device__ double direct_energy(const DirectJob &J, double x, double y, double z,
                                double4 *s_img, double *s_eps) {
    const int nb = 2 * J.rep[1] + 1, nc = 2 * J.rep[2] + 1;
    double energy = 0.0;
    for (long long q0 = 0; q0 < J.n_q; q0 += kDirectWG) {
        const long long q = q0 + threadIdx.x;
        __syncthreads();  // the previous chunk has been read by every lane
        if (q < J.n_q) {
            const long long m = q / J.n_atoms;
            const int a = (int)(q - m * J.n_atoms);
            const int i = (int)(m / (nb * nc)) - J.rep[0];
            const int j = (int)((m / nc) % nb) - J.rep[1];
            const int k = (int)(m % nc) - J.rep[2];
            const Particle A = J.atoms[a];
            double xa, ya, za;
            // src/computegridsheet.h:61-69
            frac_to_cart(J.t, A.x_f + 1.0 * i, A.y_f + 1.0 * j, A.z_f + 1.0 * k, xa, ya, za);
            s_img[threadIdx.x] = make_double4(xa, ya, za, A.sig2);
            s_eps[threadIdx.x] = A.eps;
        }
        __syncthreads();
        const int cnt = (int)(J.n_q - q0 < kDirectWG ? J.n_q - q0 : kDirectWG);
        for (int r = 0; r < cnt; ++r) {
            const double4 B = s_img[r];
            // src/computegridsheet.h:71-82
            const double dx = x - B.x, dy = y - B.y, dz = z - B.z;
            double r2 = dx * dx + dy * dy + dz * dz;
            r2 = (r2 > J.min_r2) ? r2 : J.min_r2;
            if (r2 < J.rc2) {
                const double s2 = B.w / r2;
                const double s6 = s2 * s2 * s2;
                energy += 4.0 * s_eps[r] * s6 * (s6 - 1.0);
            }
        }
    }
    return energy;
}

// Point p of the grid, index k + j*nz + i*ny*nz, at fractional
// (i/(nx-1), j/(ny-1), k/(nz-1)) (src/writegrid.cu:122-127).
__global__ __launch_bounds__(kDirectWG) void k_direct_grid(DirectJob J, int nx, int ny, int nz,
                                                           long long npts, double *out) {
    __shared__ double4 s_img[kDirectWG];
    __shared__ double s_eps[kDirectWG];
    const long long p = (long long)blockIdx.x * kDirectWG + threadIdx.x;
    const long long pp = p < npts ? p : npts - 1;
    const long long yz = (long long)ny * nz;
    const int i = (int)(pp / yz);
    const int j = (int)((pp - (long long)i * yz) / nz);
    const int k = (int)(pp - (long long)i * yz - (long long)j * nz);
    double x, y, z;
    frac_to_cart(J.t, 1.0 * i / (nx - 1), 1.0 * j / (ny - 1), 1.0 * k / (nz - 1), x, y, z);
    const double e = direct_energy(J, x, y, z, s_img, s_eps);
    if (p < npts) out[p] = e;
}

// Insertion p: the Philox4x32-10 point of the tile kernel's stream 0.
__global__ __launch_bounds__(kDirectWG) void k_direct_widom(DirectJob J, uint64_t seed,
                                                            long long nins, double *e_out) {
    __shared__ double4 s_img[kDirectWG];
    __shared__ double s_eps[kDirectWG];
    const long long p = (long long)blockIdx.x * kDirectWG + threadIdx.x;
    double f[3];
    widom_point(seed, 0u, (uint64_t)(p < nins ? p : nins - 1), f);
    double x, y, z;
    frac_to_cart(J.t, f[0], f[1], f[2], x, y, z);
    const double e = direct_energy(J, x, y, z, s_img, s_eps);
    if (p < nins) e_out[p] = e;
}

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

// Device buffers of one call, returned to the pool when the call ends (every
// path that launched work synchronises the stream first).
struct Buffers {
    Context *ctx;
    std::vector<void *> ptrs;
    template <class T>
    T *get(size_t count) {
        void *p = ctx->pool.alloc(count * sizeof(T) > 0 ? count * sizeof(T) : 16);
        if (p) ptrs.push_back(p);
        return (T *)p;
    }
    ~Buffers() {
        (void)hipStreamSynchronize(ctx->stream);
        for (void *p : ptrs) ctx->pool.free(p);
    }
};

int direct_setup(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                 const int rep[3], double rc2, double min_r2, Buffers &buf, DirectJob *J) {
    if (!ctx || (n_atoms > 0 && !atoms) || n_atoms < 0 || !t || !rep) {
        set_error("direct method: null argument or negative atom count");
        return ST_ARG;
    }
    for (int d = 0; d < 3; d++)
        if (rep[d] < 0 || rep[d] > 64) {
            set_error("direct method: replication factors must be in [0, 64]");
            return ST_ARG;
        }
    if (!(rc2 > 0.0)) {
        set_error("direct method: rc2 must be positive");
        return ST_ARG;
    }
    std::memset(J, 0, sizeof(*J));
    std::memcpy(J->t, t, sizeof(J->t));
    std::memcpy(J->rep, rep, sizeof(J->rep));
    J->n_atoms = n_atoms;
    J->n_q = (long long)n_atoms * (2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1);
    J->rc2 = rc2;
    J->min_r2 = min_r2;
    HIP_TRY(hipSetDevice(ctx->device));
    Particle *d = buf.get<Particle>(n_atoms > 0 ? n_atoms : 1);
    if (!d) return ST_NOMEM;
    if (n_atoms > 0)
        HIP_TRY(hipMemcpyAsync(d, atoms, sizeof(Particle) * n_atoms, hipMemcpyHostToDevice,
                               ctx->stream));
    J->atoms = d;
    return ST_OK;
}

}  // namespace

int direct_grid(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                const int rep[3], double rc2, const int n[3], double *host_out,
                GridDev **resident) {
    if (!n) {
        set_error("direct_grid: null argument");
        return ST_ARG;
    }
    for (int d = 0; d < 3; d++)
        if (n[d] < 2 || n[d] > 65536) {
            set_error("direct_grid: grid dimensions must be in [2, 65536]");
            return ST_ARG;
        }
    const long long npts = (long long)n[0] * n[1] * n[2];
    if (npts > (1LL << 31)) {
        set_error("direct_grid: at most 2^31 points");
        return ST_ARG;
    }
    if (!ctx) {
        set_error("direct_grid: null context");
        return ST_ARG;
    }
    Buffers buf{ctx, {}};
    DirectJob J;
    const double min_r = 1e-12;  // src/computegridsheet.h:19
    int st = direct_setup(ctx, atoms, n_atoms, t, rep, rc2, min_r * min_r, buf, &J);
    if (st != ST_OK) return st;
    GridDev *g = new GridDev();
    g->ctx = ctx;
    for (int d = 0; d < 3; d++) g->n[d] = n[d];
    g->precision = 1;
    g->d_data = ctx->pool.alloc(sizeof(double) * npts);
    if (!g->d_data) {
        delete g;
        return ST_NOMEM;
    }
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_direct_grid, dim3((unsigned)((npts + kDirectWG - 1) / kDirectWG)),
                       dim3(kDirectWG), 0, s, J, n[0], n[1], n[2], npts, (double *)g->d_data);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && host_out)
        e = hipMemcpyAsync(host_out, g->d_data, sizeof(double) * npts, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        delete g;
        set_error(std::string("direct_grid: ") + hipGetErrorName(e));
        return ST_HIP;
    }
    if (resident) *resident = g;
    else delete g;
    return ST_OK;
}

int direct_widom(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                 const int rep[3], double rc2, uint64_t n_ins, double temperature, uint64_t seed,
                 WidomResult *out, double *energies_host) {
    if (!ctx || !out) {
        set_error("direct_widom: null argument");
        return ST_ARG;
    }
    if (n_ins < 1 || n_ins > (1ULL << 31) || !(temperature > 0.0)) {
        set_error("direct_widom: insertions in [1, 2^31] and a positive temperature");
        return ST_ARG;
    }
    Buffers buf{ctx, {}};
    DirectJob J;
    const double min_r = 1e-16;  // src/doHenry.h:15
    int st = direct_setup(ctx, atoms, n_atoms, t, rep, rc2, min_r * min_r, buf, &J);
    if (st != ST_OK) return st;
    double *e = buf.get<double>(n_ins);
    if (!e) return ST_NOMEM;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_direct_widom, dim3((unsigned)((n_ins + kDirectWG - 1) / kDirectWG)),
                       dim3(kDirectWG), 0, s, J, seed, (long long)n_ins, e);
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_widom_partials(e, (long long)n_ins, temperature, ctx->d_partials, s));
    HIP_TRY(launch_widom_final(ctx->d_partials, ctx->d_sums, s));
    double sums[2];
    HIP_TRY(hipMemcpyAsync(sums, ctx->d_sums, sizeof(sums), hipMemcpyDeviceToHost, s));
    if (energies_host)
        HIP_TRY(hipMemcpyAsync(energies_host, e, sizeof(double) * n_ins, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    out->sum_b = sums[0];
    out->sum_eb = sums[1];
    out->n = n_ins;
    return ST_OK;
}

}  // namespace er
