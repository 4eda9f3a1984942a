// er_kernels.hip -- gfx950 kernels for the guest-framework Lennard-Jones
// energy grid (src/computegridsheet.h:30-99) and Widom insertion
// (src/doHenry.h:29-135).
//
// Layout in HBM, per framework:
//   atoms   : AtomDev[n]           fractional coordinates + type, grouped by type
//   near    : double4[]            Cartesian FP64 image atoms T(f_a + n) that can
//                                  lie within r_c of the unit cell, grouped by type,
//                                  reference enumeration order inside a type
//   types   : TypeParams[ntypes]   mixed sigma^2, 4 eps, FP64 switch radius
// The tiled kernel (k_tiles) is persistent: each workgroup takes work items
// (a region of grid tiles or of Widom bins) from a device counter, gathers the
// near-list indices of the region's candidates into its scratch slice, then
// per tile (a box of grid points, or one bin's insertions) keeps the image
// atoms within r_c + R_tile in LDS, split into far records (inside the cutoff
// and beyond the FP64 switch radius for every point: no test, FP32
// reciprocal powers), shell records (cutoff test) and near records (FP64
// pair arithmetic).  Every lane evaluates its points against the LDS list
// with the record broadcast from LDS and the pair parameters in registers;
// in FP32 and FP64 grid tiles each wave first re-sorts the shell records
// against its own, smaller point box.  A Widom bin of at most one chunk of
// insertions keeps their points and partial energies in registers across
// its LDS passes.
#include "er_kernels.h"

#include <float.h>

#include <type_traits>

// No implicit multiply-add contraction in this file: every FMA below is
// written out, so the arithmetic of a pair (and so every energy) is fixed by
// the source and does not depend on how the compiler schedules one template
// instance against another (1, 2 or 4 points per lane give bit-identical
// energies for the same list).  Coordinates (frac_to_cart) then round like
// the oracle's and the reference's host code: separate products and sums.
#pragma clang fp contract(off)

namespace er {

namespace {

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ void frac_to_cart(const double *t, double fx, double fy, double fz,
                                             double &x, double &y, double &z) {
    x = t[0] * fx + t[1] * fy + t[2] * fz;
    y = t[3] * fx + t[4] * fy + t[5] * fz;
    z = t[6] * fx + t[7] * fy + t[8] * fz;
}

// ---------------------------------------------------------------------------
// Near-image list
// ---------------------------------------------------------------------------

struct NearItem {
    bool valid;
    double gx, gy, gz;  // fractional position of the image atom
};

__device__ __forceinline__ NearItem near_item(const NearPlan &p, int blk, int tid) {
    NearItem it;
    it.valid = false;
    int t = 0;
    while (t + 1 < p.ntypes && p.type_block0[t + 1] <= blk) ++t;
    const int a0 = p.type_atom0[t];
    const int cnt = p.type_atom0[t + 1] - a0;
    const long long q = (long long)(blk - p.type_block0[t]) * kWG + tid;
    if (cnt <= 0 || q >= (long long)cnt * p.n_images) return it;
    const int m = (int)(q / cnt);
    const int a = a0 + (int)(q % cnt);
    const int nb = 2 * p.rep[1] + 1, nc = 2 * p.rep[2] + 1;
    const int ii = m / (nb * nc) - p.rep[0];
    const int jj = (m / nc) % nb - p.rep[1];
    const int kk = m % nc - p.rep[2];
    const AtomDev at = p.atoms[a];
    // src/computegridsheet.h:61-63
    it.gx = at.fx + 1.0 * ii;
    it.gy = at.fy + 1.0 * jj;
    it.gz = at.fz + 1.0 * kk;
    it.valid = it.gx >= -p.ext[0] && it.gx <= 1.0 + p.ext[0] && it.gy >= -p.ext[1] &&
               it.gy <= 1.0 + p.ext[1] && it.gz >= -p.ext[2] && it.gz <= 1.0 + p.ext[2];
    return it;
}

__global__ __launch_bounds__(kWG) void k_near_count(NearPlan p, uint32_t *block_count) {
    __shared__ int s_w[kWG / 64];
    const int tid = threadIdx.x;
    NearItem it = near_item(p, blockIdx.x, tid);
    uint64_t m = __ballot(it.valid);
    if ((tid & 63) == 0) s_w[tid >> 6] = __popcll(m);
    __syncthreads();
    if (tid == 0) block_count[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(kWG) void k_near_write(NearPlan p, const uint32_t *block_off,
                                                    double4 *near) {
    __shared__ int s_w[kWG / 64];
    const int tid = threadIdx.x, wave = tid >> 6;
    NearItem it = near_item(p, blockIdx.x, tid);
    uint64_t m = __ballot(it.valid);
    if ((tid & 63) == 0) s_w[wave] = __popcll(m);
    __syncthreads();
    if (it.valid) {
        int off = block_off[blockIdx.x] + lanes_below(m);
        for (int w = 0; w < wave; ++w) off += s_w[w];
        double x, y, z;
        frac_to_cart(p.t, it.gx, it.gy, it.gz, x, y, z);
        near[off] = make_double4(x, y, z, 0.0);
    }
}

// Single-workgroup exclusive scan of n uint32 values; out[n] = total.
__device__ void block_exclusive_scan(const uint32_t *in, long long n, uint32_t *out,
                                     uint32_t *out2) {
    __shared__ uint32_t s_part[kWG];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    const int per = 8;
    for (long long base = 0; base < n; base += (long long)kWG * per) {
        uint32_t v[per];
        uint32_t sum = 0;
        for (int k = 0; k < per; ++k) {
            long long i = base + (long long)tid * per + k;
            v[k] = (i < n) ? in[i] : 0u;
            sum += v[k];
        }
        s_part[tid] = sum;
        __syncthreads();
        // Hillis-Steele inclusive scan over the 256 partial sums
        for (int d = 1; d < kWG; d <<= 1) {
            uint32_t add = (tid >= d) ? s_part[tid - d] : 0u;
            __syncthreads();
            s_part[tid] += add;
            __syncthreads();
        }
        uint32_t run = s_carry + s_part[tid] - sum;
        for (int k = 0; k < per; ++k) {
            long long i = base + (long long)tid * per + k;
            if (i < n) {
                out[i] = run;
                if (out2) out2[i] = run;
            }
            run += v[k];
        }
        __syncthreads();
        if (tid == kWG - 1) s_carry += s_part[kWG - 1];
        __syncthreads();
    }
    if (tid == 0) {
        out[n] = s_carry;
        if (out2) out2[n] = s_carry;
    }
}

__global__ __launch_bounds__(kWG) void k_near_offsets(NearPlan p, const uint32_t *block_count,
                                                      uint32_t *block_off, int *type_begin) {
    block_exclusive_scan(block_count, p.nblocks, block_off, nullptr);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 0; t < p.ntypes; ++t) type_begin[t] = (int)block_off[p.type_block0[t]];
        type_begin[p.ntypes] = (int)block_off[p.nblocks];
    }
}

// ---------------------------------------------------------------------------
// Widom binning
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t widom_bin(const double f[3], const int nb[3]) {
    int b[3];
    for (int d = 0; d < 3; ++d) {
        int v = (int)(f[d] * nb[d]);  // f in (0,1]
        b[d] = v < nb[d] ? v : nb[d] - 1;
    }
    return (uint32_t)((b[0] * nb[1] + b[1]) * nb[2] + b[2]);
}

// Batched binning.  Each block takes `chunk` consecutive insertion indices of
// one job, histograms their bins in LDS and touches global memory once per
// bin it saw: the count pass adds its counts, the scatter pass reserves a
// range of the bin's slots and hands them out through LDS atomics.  The order
// inside a bin is therefore arbitrary; no result depends on it (every
// insertion of a bin sees the same candidate list, and sums run over
// insertion indices).  The histogram lives in dynamic LDS sized per launch
// for the largest job's bins (lds_bins, at most kLdsBinsMax); jobs with more
// bins than that use one global atomic per insertion.
__device__ __forceinline__ int bin_job_of(const BinJob *__restrict__ jobs, int njobs,
                                          long long blk) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].block0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(kWG) void k_bin_zero(const BinJob *__restrict__ jobs) {
    const BinJob &J = jobs[blockIdx.x];
    const long long nbt = (long long)J.nb[0] * J.nb[1] * J.nb[2];
    for (long long b = threadIdx.x; b <= nbt; b += kWG) J.count[b] = 0u;
}

__global__ __launch_bounds__(kWG) void k_bin_count(const BinJob *__restrict__ jobs, int njobs,
                                                   int chunk, int lds_bins) {
    extern __shared__ uint32_t s_hist[];
    const int j = bin_job_of(jobs, njobs, blockIdx.x);
    const BinJob &J = jobs[j];
    const uint64_t i0 = (uint64_t)(blockIdx.x - J.block0) * (uint64_t)chunk;
    const uint64_t i1 = i0 + (uint64_t)chunk < J.n ? i0 + (uint64_t)chunk : J.n;
    const int nbt = J.nb[0] * J.nb[1] * J.nb[2];
    if (nbt > lds_bins) {
        for (uint64_t i = i0 + threadIdx.x; i < i1; i += kWG) {
            double f[3];
            widom_point(J.seed, J.stream, i, f);
            atomicAdd(&J.count[widom_bin(f, J.nb)], 1u);
        }
        return;
    }
    for (int b = threadIdx.x; b < nbt; b += kWG) s_hist[b] = 0u;
    __syncthreads();
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kWG) {
        double f[3];
        widom_point(J.seed, J.stream, i, f);
        atomicAdd(&s_hist[widom_bin(f, J.nb)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbt; b += kWG)
        if (s_hist[b]) atomicAdd(&J.count[b], s_hist[b]);
}

__global__ __launch_bounds__(kWG) void k_bin_scan(const BinJob *__restrict__ jobs) {
    const BinJob &J = jobs[blockIdx.x];
    const long long nbt = (long long)J.nb[0] * J.nb[1] * J.nb[2];
    block_exclusive_scan(J.count, nbt, J.start, J.cursor);
}

__global__ __launch_bounds__(kWG) void k_bin_scatter(const BinJob *__restrict__ jobs, int njobs,
                                                     int chunk, int lds_bins) {
    extern __shared__ uint32_t s_hist[];
    const int j = bin_job_of(jobs, njobs, blockIdx.x);
    const BinJob &J = jobs[j];
    const uint64_t i0 = (uint64_t)(blockIdx.x - J.block0) * (uint64_t)chunk;
    const uint64_t i1 = i0 + (uint64_t)chunk < J.n ? i0 + (uint64_t)chunk : J.n;
    const int nbt = J.nb[0] * J.nb[1] * J.nb[2];
    if (nbt > lds_bins) {
        for (uint64_t i = i0 + threadIdx.x; i < i1; i += kWG) {
            double f[3];
            widom_point(J.seed, J.stream, i, f);
            const uint32_t pos = atomicAdd(&J.cursor[widom_bin(f, J.nb)], 1u);
            J.sorted[pos] = (uint32_t)i;
        }
        return;
    }
    for (int b = threadIdx.x; b < nbt; b += kWG) s_hist[b] = 0u;
    __syncthreads();
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kWG) {
        double f[3];
        widom_point(J.seed, J.stream, i, f);
        atomicAdd(&s_hist[widom_bin(f, J.nb)], 1u);
    }
    __syncthreads();
    // reserve each bin's share of slots; s_hist then holds this block's next slot
    for (int b = threadIdx.x; b < nbt; b += kWG) {
        const uint32_t h = s_hist[b];
        if (h) s_hist[b] = atomicAdd(&J.cursor[b], h);
    }
    __syncthreads();
    for (uint64_t i = i0 + threadIdx.x; i < i1; i += kWG) {
        double f[3];
        widom_point(J.seed, J.stream, i, f);
        const uint32_t pos = atomicAdd(&s_hist[widom_bin(f, J.nb)], 1u);
        J.sorted[pos] = (uint32_t)i;
    }
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Tiled energy kernel
// ---------------------------------------------------------------------------

// Margin (Angstrom) applied to every list classification so that FP64
// rounding of the tile radius and distances can never misclassify a pair.
constexpr double kMargin = 1e-6;

enum : int { CAT_NONE = 0, CAT_FAR = 1, CAT_SHELL = 2, CAT_NEAR = 3 };

// Ints per list segment: type, far begin/end, shell begin/end, near begin/end.
constexpr int kSegInts = 7;

// Pairs per point summed in FP32 (far and shell pairs of one type) before the
// sum is scaled and added to the FP64 energy.
constexpr int kFlushPairs = 16;

struct Tile {
    double cx, cy, cz, R;
};

__device__ __forceinline__ Tile tile_sphere(const double *t, const double lo[3],
                                            const double hi[3]) {
    double mid[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    Tile s;
    frac_to_cart(t, mid[0], mid[1], mid[2], s.cx, s.cy, s.cz);
    double r2 = 0.0;
    for (int c = 0; c < 8; ++c) {
        double d[3] = {((c & 4) ? hi[0] : lo[0]) - mid[0], ((c & 2) ? hi[1] : lo[1]) - mid[1],
                       ((c & 1) ? hi[2] : lo[2]) - mid[2]};
        double x, y, z;
        frac_to_cart(t, d[0], d[1], d[2], x, y, z);
        double q = x * x + y * y + z * z;
        r2 = q > r2 ? q : r2;
    }
    s.R = sqrt(r2) * (1.0 + 1e-12) + kMargin;
    return s;
}

// Split of the LDS list: [0, capA) holds far records (from 0 upwards) and
// shell records (from capA-1 downwards); [capA, cap) holds near records.
template <int MATH>
__device__ __forceinline__ int near_region_start(int cap) {
    if (MATH == MATH_FP64) return cap - kWG;  // near: atoms within R of the tile centre
    if (MATH == MATH_COUNT || MATH == MATH_OVERLAP) return cap - kWG;  // no near records
    return mixed_far_shell_cap(cap);
}

// Stream the near list of job J from (cur_t, cur_pos) into the LDS list until
// it is exhausted or a region is full, classifying each image atom by its
// distance d to the tile centre: far (every point of the tile is inside the
// cutoff and beyond the FP64 switch radius), shell (beyond the switch radius
// but possibly outside the cutoff), near (possibly inside the switch radius).
// Records hold the atom relative to the tile centre and its squared distance,
// (a - c, |a - c|^2), and keep the source order, so the list (and every sum
// over it) depends only on the tile and the source.  The source is the near
// list src, position p holding src[p], or, with sidx, a sub-list of it,
// position p holding src[sidx[p]].  Returns the number of segments written
// to seg.
template <int MATH>
__device__ int fill_list(const JobDev *__restrict__ J, const double4 *__restrict__ src,
                         const uint32_t *__restrict__ sidx, const int *__restrict__ tbeg,
                         const Tile &tl, int &cur_t, int &cur_pos,
                         int cap, double4 *__restrict__ L, int *__restrict__ seg,
                         int *__restrict__ s_cnt) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntypes = J->ntypes;
    const double rc = J->rc;
    const int capA = near_region_start<MATH>(cap);
    const int capN = cap - capA;
    // Squared-distance forms of the tests on d = |image atom - tile centre|:
    //   some point may be inside r_c:     d - R <  r_c             (in2)
    //   every point is inside r_c:        d + R <  r_c             (far2)
    //   some point may be inside r_FP64:  d - R <= r_FP64          (near2, per type)
    // each widened by kMargin so FP64 rounding cannot misclassify.
    const double in2 = (rc + tl.R + kMargin) * (rc + tl.R + kMargin);
    const double far_r = rc - tl.R - kMargin;
    const double far2 = far_r > 0.0 ? far_r * far_r : -1.0;
    // FP64 grids: near (clamp possible) only if some point may lie within
    // kMargin of the atom, d <= R + kMargin
    const double clamp2 = (tl.R + kMargin) * (tl.R + kMargin);
    // The cell's points lie in a parallelepiped around the centre:
    // J->cell_box[d] times cell vector t_d on either side (a bin's box; a
    // full tile's, which holds a partial edge tile's around that tile's own
    // centre).  A shell record at v from the centre with |v|^2 - h(v) >
    // (r_c + m)|v|, h(v) = sum_d cell_box_d |t_d . v|, is beyond r_c of every
    // point (|a - x| >= |v| - h(v)/|v|) and adds exactly 0: it is not kept.
    // (FP32 and FP64 grid tiles re-test their shell records per wave, against
    // the wave's smaller box, which makes this test redundant there)
    const bool wave_lists = J->kind == JOB_GRID && (MATH == MATH_MIXED || MATH == MATH_FP64);
    const double *hb = J->cell_box;
    const double rout = rc + kMargin, rout2 = rout * rout;
    const double *tm = J->t;
    double near2 = -1.0;
    int near_t = -1;
    int nf = 0, ns = 0, nn = 0, nseg = 0, f0 = 0, s0 = 0, n0 = 0;
    while (cur_t < ntypes) {
        if (MATH == MATH_MIXED && near_t != cur_t) {
            const double rn = J->types[cur_t].rclose + tl.R + kMargin;
            near2 = rn * rn;
            near_t = cur_t;
        }
        const int tend = uniform(tbeg[cur_t + 1]);
        if (cur_pos >= tend) {
            if (nf > f0 || ns > s0 || nn > n0) {
                if (tid == 0) {
                    int *g = seg + kSegInts * nseg;
                    g[0] = cur_t; g[1] = f0; g[2] = nf; g[3] = s0; g[4] = ns; g[5] = n0; g[6] = nn;
                }
                ++nseg;
            }
            f0 = nf;
            s0 = ns;
            n0 = nn;
            ++cur_t;
            if (cur_t < ntypes) cur_pos = uniform(tbeg[cur_t]);
            continue;
        }
        if (nseg >= kMaxSeg) break;
        const int idx = cur_pos + tid;
        int cat = CAT_NONE;
        double4 A = make_double4(0.0, 0.0, 0.0, 0.0);
        if (idx < tend) {
            const double4 S = src[sidx ? (int)sidx[idx] : idx];
            const double dx = S.x - tl.cx, dy = S.y - tl.cy, dz = S.z - tl.cz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            A = make_double4(dx, dy, dz, d2);
            if (d2 < in2) {
                if (MATH == MATH_FP64 && d2 <= clamp2) cat = CAT_NEAR;
                else if (MATH == MATH_MIXED && d2 <= near2) cat = CAT_NEAR;
                else cat = (MATH != MATH_OVERLAP && d2 < far2) ? CAT_FAR : CAT_SHELL;
                if (cat == CAT_SHELL && !wave_lists) {
                    const double h = hb[0] * fabs(tm[0] * dx + tm[3] * dy + tm[6] * dz) +
                                     hb[1] * fabs(tm[1] * dx + tm[4] * dy + tm[7] * dz) +
                                     hb[2] * fabs(tm[2] * dx + tm[5] * dy + tm[8] * dz);
                    const double g = d2 - h;
                    if (g > 0.0 && g * g > rout2 * d2) cat = CAT_NONE;
                }
            }
        }
        const uint64_t bf = __ballot(cat == CAT_FAR);
        const uint64_t bs = __ballot(cat == CAT_SHELL);
        const uint64_t bn = __ballot(cat == CAT_NEAR);
        if (lane == 0) {
            s_cnt[3 * wave] = __popcll(bf);
            s_cnt[3 * wave + 1] = __popcll(bs);
            s_cnt[3 * wave + 2] = __popcll(bn);
        }
        __syncthreads();
        int of = nf, os = ns, on = nn, tf = 0, ts = 0, tn = 0;
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) {
            const int cf = s_cnt[3 * w], cs = s_cnt[3 * w + 1], cn = s_cnt[3 * w + 2];
            if (w < wave) {
                of += cf;
                os += cs;
                on += cn;
            }
            tf += cf;
            ts += cs;
            tn += cn;
        }
        // A pass takes the kept records in source order up to the first one
        // that does not fit (far + shell share [0, capA), near [capA, cap)),
        // so its content depends only on the sequence of kept records, not on
        // the source's chunking: a region sub-list and the near list give the
        // same passes.  Records after that one may also be written (into
        // slots past the final counts, never read in this pass).
        const int fs_before = of + os + lanes_below(bf | bs);
        const int n_before = on + lanes_below(bn);
        const bool fits = (cat == CAT_FAR || cat == CAT_SHELL) ? fs_before < capA
                          : cat == CAT_NEAR                    ? n_before < capN
                                                               : true;
        if (fits && cat == CAT_FAR) L[of + lanes_below(bf)] = A;
        if (fits && cat == CAT_SHELL) L[capA - 1 - (os + lanes_below(bs))] = A;
        if (fits && cat == CAT_NEAR) L[capA + on + lanes_below(bn)] = A;  // A: (a - c, |a - c|^2)
        const uint64_t bx = __ballot(!fits);
        if (lane == 0) {
            // this wave's first record that does not fit, and the counts before it
            int *x = s_cnt + 3 * (kWG / 64) + 4 * wave;
            if (bx) {
                const int l = __builtin_ctzll(bx);
                const uint64_t below = (1ull << l) - 1ull;
                x[0] = l;
                x[1] = of + __popcll(bf & below);
                x[2] = os + __popcll(bs & below);
                x[3] = on + __popcll(bn & below);
            } else {
                x[0] = 64;
            }
        }
        __syncthreads();
        int stop = -1;
#pragma unroll
        for (int w = kWG / 64 - 1; w >= 0; --w)
            if (s_cnt[3 * (kWG / 64) + 4 * w] < 64) stop = w;
        if (stop >= 0) {
            const int *x = s_cnt + 3 * (kWG / 64) + 4 * stop;
            nf = uniform(x[1]);
            ns = uniform(x[2]);
            nn = uniform(x[3]);
            cur_pos += 64 * stop + uniform(x[0]);
            __syncthreads();  // x is read by every thread before it is reused
            break;
        }
        nf = uniform(nf + tf);
        ns = uniform(ns + ts);
        nn = uniform(nn + tn);
        cur_pos += kWG;
    }
    if (nf > f0 || ns > s0 || nn > n0) {
        if (tid == 0) {
            int *g = seg + kSegInts * nseg;
            g[0] = cur_t; g[1] = f0; g[2] = nf; g[3] = s0; g[4] = ns; g[5] = n0; g[6] = nn;
        }
        ++nseg;
    }
    return nseg;
}

// r^2 from this lane's points to record A (both relative to the tile centre):
// the expanded form |a|^2 + |p|^2 - 2 p.a with the z, y, x terms in that
// order.  With XL the lane's points share y and z (grid points that differ
// only in i, on a cell whose first vector lies along x), so the z and y terms
// are formed once per record; the result is bit-identical to the per-point
// form, since the shared terms are the same operations on the same values.
template <int PPL, bool XL>
__device__ __forceinline__ void pair_r2(const double4 &A, const double (&mx)[PPL],
                                        const double (&my)[PPL], const double (&mz)[PPL],
                                        const double (&p2)[PPL], double (&r2)[PPL]) {
    if (XL) {
        const double yz = fma(my[0], A.y, fma(mz[0], A.z, A.w));
#pragma unroll
        for (int q = 0; q < PPL; ++q) r2[q] = fma(mx[q], A.x, yz) + p2[q];
    } else {
#pragma unroll
        for (int q = 0; q < PPL; ++q)
            r2[q] = fma(mx[q], A.x, fma(my[q], A.y, fma(mz[q], A.z, A.w))) + p2[q];
    }
}

// Shell records [sb, se) of a segment (record k at Ls[-k]) against this
// wave's sphere wsph (centre relative to the tile centre, radius wsph.w):
// records beyond r_c + w of it are out of the cutoff for all of the wave's
// points and dropped, those within r_c - w are inside it for all of them (no
// cutoff test).  The wave's points lie in a parallelepiped around that
// centre, J->wave_box[d] times cell vector t_d on either side; for a record
// at v from the centre every point x has |a - x| >= (a - x).v/|v| >=
// |v| - h(v)/|v|, h(v) = sum_d wave_box_d |t_d . v|, so a record with
// |v|^2 - h(v) > (r_c + m)|v| is out of the cutoff for every point as well
// and is dropped: it added exactly 0 to every sum.  (The matching bound
// |a - x|^2 <= |v|^2 + 2 h(v) + w^2 would move a few records from the
// shell to the far list; it measured no gain.)  The wave writes its own
// list of LDS byte offsets (from the start of the list, capA - 1 - k
// records), order kept: the nwf wave-far ones from the front of wl, the nws
// still-shell ones from wl[wcap - 1] down (se - sb <= wcap entries).
__device__ __forceinline__ void wave_shell_lists(const JobDev *__restrict__ J,
                                                 const double4 *__restrict__ Ls, int capA,
                                                 int wcap, int sb, int se, const double4 &wsph,
                                                 uint16_t *__restrict__ wl, int &nwf, int &nws) {
    const int lane = threadIdx.x & 63;
    const double rin = J->rc + wsph.w + kMargin, rfar = J->rc - wsph.w - kMargin;
    const double in2 = rin * rin, far2 = rfar > 0.0 ? rfar * rfar : -1.0;
    const double rout = J->rc + kMargin, rout2 = rout * rout;
    const double *t = J->t, *wb = J->wave_box;
    nwf = 0;
    nws = 0;
    for (int k0 = sb; k0 < se; k0 += 64) {
        const int k = k0 + lane;
        int cat = 0;
        if (k < se) {
            const double4 A = Ls[-k];
            const double dx = A.x - wsph.x, dy = A.y - wsph.y, dz = A.z - wsph.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            cat = d2 < in2 ? (d2 < far2 ? 1 : 2) : 0;
            if (cat == 2) {
                const double h = wb[0] * fabs(t[0] * dx + t[3] * dy + t[6] * dz) +
                                 wb[1] * fabs(t[1] * dx + t[4] * dy + t[7] * dz) +
                                 wb[2] * fabs(t[2] * dx + t[5] * dy + t[8] * dz);
                const double g = d2 - h;
                if (g > 0.0 && g * g > rout2 * d2) cat = 0;
            }
        }
        const uint64_t bf = __ballot(cat == 1), bs = __ballot(cat == 2);
        const uint16_t off = (uint16_t)((capA - 1 - k) * (int)sizeof(double4));
        if (cat == 1) wl[nwf + lanes_below(bf)] = off;
        if (cat == 2) wl[wcap - 1 - (nws + lanes_below(bs))] = off;
        nwf = uniform(nwf + __popcll(bf));
        nws = uniform(nws + __popcll(bs));
    }
    // the lists are read back by every lane of this wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Energies (or in-cutoff counts) of this lane's PPL points against the LDS
// list.  Points and records are relative to the tile centre c.
//   Far pairs (inside r_c and beyond the FP64 switch radius for every point
//   of the tile): r^2 = |a|^2 + |p|^2 - 2 p.a in FP64 (three FMAs and an add;
//   |a| <= r_c + R keeps the cancellation below 13x), then in FP32
//   u = 1/r^2, and per point the sums S3 = sum u^3 and S6 = sum u^6 over at
//   most kFlushPairs pairs of one type, flushed into the FP64 energy as
//   4 eps (sigma^12 S6 - sigma^6 S3).  sigma and eps stay FP64, so their
//   rounding adds no error common to every pair of a type.
//   Shell pairs: the same, behind the strict FP64 test r^2 < r_c^2.
//   Near pairs (every pair of an FP64 grid): the same FP64 r^2, the
//   reference's clamp and strict cutoff test, FP64 LJ with a Newton-refined
//   reciprocal (one step for FP32 grids and Widom, two for FP64 grids).
// XL: see pair_r2.
// ws (MIXED and FP64 grid tiles, wave-uniform): this wave's points lie in the
// sphere wsph (centre relative to the tile centre, radius w), smaller than
// the tile's, and in the parallelepiped of J->wave_box around the same
// centre; each segment's shell records are then re-classified against them
// into the wave's own lists of LDS byte offsets wl (wave_shell_lists).
// phase (wave-uniform): 0 every record, per segment far, shell, near; 1 only
// the near records; 2 only the far and shell records.
template <int PPL, int MATH, bool XL>
__device__ __forceinline__ void eval_list(const JobDev *__restrict__ J,
                                          const double4 *__restrict__ L, int cap,
                                          const int *__restrict__ seg, int nseg,
                                          const double (&px)[PPL], const double (&py)[PPL],
                                          const double (&pz)[PPL], double (&acc)[PPL],
                                          double R, bool ws = false,
                                          double4 wsph = make_double4(0.0, 0.0, 0.0, 0.0),
                                          uint16_t *__restrict__ wl = nullptr, int phase = 0) {
    const bool do_fs = phase != 1, do_near = phase != 2;
    const double rc2 = J->rc2, min_r2 = J->min_r2;
    const int capA = near_region_start<MATH>(cap);
    const double4 *__restrict__ Ls = L + capA - 1;  // shell k at Ls[-k]
    const double4 *__restrict__ Ln = L + capA;      // near k at Ln[k]
    double mx[PPL], my[PPL], mz[PPL], p2[PPL];
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        mx[q] = -2.0 * px[q];
        my[q] = -2.0 * py[q];
        mz[q] = -2.0 * pz[q];
        p2[q] = px[q] * px[q] + py[q] * py[q] + pz[q] * pz[q];
    }
    for (int s = 0; s < nseg; ++s) {
        const int t = uniform(seg[kSegInts * s]);
        const int fb = uniform(seg[kSegInts * s + 1]), fe = uniform(seg[kSegInts * s + 2]);
        const int sb = uniform(seg[kSegInts * s + 3]), se = uniform(seg[kSegInts * s + 4]);
        const int nb = uniform(seg[kSegInts * s + 5]), ne = uniform(seg[kSegInts * s + 6]);
        const TypeParams P = J->types[t];
        if (MATH == MATH_OVERLAP) {
            // src/computegridsheet.h:154: overlap iff r^2 < sigma_ij^2 (every
            // record is a shell record: the cutoff is sigma_max)
            const double sig2 = P.sig2;
            for (int k = sb; k < se; ++k) {
                double r2[PPL];
                pair_r2<PPL, XL>(Ls[-k], mx, my, mz, p2, r2);
#pragma unroll
                for (int q = 0; q < PPL; ++q) acc[q] += r2[q] < sig2 ? 1.0 : 0.0;
            }
            continue;
        }
        if (MATH == MATH_COUNT) {
            if (!do_fs) continue;  // counting lists hold no near records
#pragma unroll
            for (int q = 0; q < PPL; ++q) acc[q] += (double)(fe - fb);
            for (int k = sb; k < se; ++k) {
                double r2[PPL];
                pair_r2<PPL, XL>(Ls[-k], mx, my, mz, p2, r2);
#pragma unroll
                for (int q = 0; q < PPL; ++q) acc[q] += r2[q] < rc2 ? 1.0 : 0.0;
            }
            continue;
        }
        if (MATH == MATH_MIXED) {
            const double sig6 = P.sig2 * P.sig2 * P.sig2;
            const double c6 = P.eps4 * sig6 * sig6, c3 = -P.eps4 * sig6;
            auto far_pair = [&](const double4 &A, float (&s3)[PPL], float (&s6)[PPL]) {
                double r2[PPL];
                pair_r2<PPL, XL>(A, mx, my, mz, p2, r2);
#pragma unroll
                for (int q = 0; q < PPL; ++q) {
                    const float w = __builtin_amdgcn_rcpf((float)r2[q]);
                    const float w2 = w * w, w3 = w2 * w;
                    s3[q] = __builtin_fmaf(w2, w, s3[q]);
                    s6[q] = __builtin_fmaf(w3, w3, s6[q]);
                }
            };
            auto shell_pair = [&](const double4 &A, float (&s3)[PPL], float (&s6)[PPL]) {
                double r2[PPL];
                pair_r2<PPL, XL>(A, mx, my, mz, p2, r2);
#pragma unroll
                for (int q = 0; q < PPL; ++q) {
                    const float w = __builtin_amdgcn_rcpf((float)r2[q]);
                    float w3 = w * w * w;
                    w3 = r2[q] < rc2 ? w3 : 0.0f;  // src/computegridsheet.h:79
                    s3[q] += w3;
                    s6[q] = __builtin_fmaf(w3, w3, s6[q]);
                }
            };
            // far records [fb, fe), then shell records [sb, se): FP32 sums of
            // kFlushPairs pairs per point (blocks of 8 unrolled records),
            // then flushed into the FP64 energy
            auto blocks = [&](int b, int e, auto &&pair_at) {
                for (int r = b; r < e; r += kFlushPairs) {
                    float a3[PPL], a6[PPL];
#pragma unroll
                    for (int q = 0; q < PPL; ++q) a3[q] = a6[q] = 0.0f;
                    const int re = r + kFlushPairs < e ? r + kFlushPairs : e;
#pragma unroll 1
                    for (int h = r; h < re; h += 8) {
                        if (h + 8 <= re) {
#pragma unroll
                            for (int u = 0; u < 8; ++u) pair_at(h + u, a3, a6);
                        } else {
                            for (int k = h; k < re; ++k) pair_at(k, a3, a6);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < PPL; ++q)
                        acc[q] = fma(c6, (double)a6[q], fma(c3, (double)a3[q], acc[q]));
                }
            };
            if (do_fs)
                blocks(fb, fe, [&](int k, float (&s3)[PPL], float (&s6)[PPL]) {
                    far_pair(L[k], s3, s6);
                });
            if (!do_fs) {
            } else if (ws) {
                int nwf, nws;
                wave_shell_lists(J, Ls, capA, capA, sb, se, wsph, wl, nwf, nws);
                const char *__restrict__ Lb = reinterpret_cast<const char *>(L);
                auto rec = [&](int e) {
                    return *reinterpret_cast<const double4 *>(Lb + wl[e]);
                };
                blocks(0, nwf, [&](int e, float (&s3)[PPL], float (&s6)[PPL]) {
                    far_pair(rec(e), s3, s6);
                });
                blocks(0, nws, [&](int e, float (&s3)[PPL], float (&s6)[PPL]) {
                    shell_pair(rec(capA - 1 - e), s3, s6);
                });
            } else {
                blocks(sb, se, [&](int k, float (&s3)[PPL], float (&s6)[PPL]) {
                    shell_pair(Ls[-k], s3, s6);
                });
            }
            if (!do_near) continue;
            // Near pairs of an FP32 grid or Widom run: the expanded FP64 r^2 of
            // the far pairs (|a| <= r_FP64 + R keeps it exact to ~1e-14 A^2),
            // the reference's clamp and strict cutoff, u = 1/r^2 from the FP32
            // reciprocal and one Newton step (~1e-14 relative), and FP64 sums
            // of u^3 and u^6 scaled by 4 eps sigma^6 / sigma^12 per segment.
            double b3[PPL], b6[PPL];
#pragma unroll
            for (int q = 0; q < PPL; ++q) b3[q] = b6[q] = 0.0;
            // A near record lies within r_FP64 + R of the tile centre, so
            // within r_FP64 + 2R of every point: when that is inside r_c the
            // strict cutoff test (src/computegridsheet.h:79) passes for every
            // pair and is skipped.
            const bool all_in = P.rclose + 2.0 * R + 3.0 * kMargin < J->rc;
            auto near_pairs = [&](auto test) {
                for (int k = nb; k < ne; ++k) {
                    double r2[PPL];
                    pair_r2<PPL, XL>(Ln[k], mx, my, mz, p2, r2);
#pragma unroll
                    for (int q = 0; q < PPL; ++q) {
                        // src/computegridsheet.h:77 (r^2 is never NaN, so fmax is the clamp)
                        const double rr = fmax(r2[q], min_r2);
                        double y = (double)__builtin_amdgcn_rcpf((float)rr);
                        y = fma(fma(-rr, y, 1.0), y, y);
                        double y3 = y * y * y;
                        if (decltype(test)::value) y3 = rr < rc2 ? y3 : 0.0;  // :79
                        b3[q] += y3;
                        b6[q] = fma(y3, y3, b6[q]);
                    }
                }
            };
            if (all_in) near_pairs(std::integral_constant<bool, false>());
            else near_pairs(std::integral_constant<bool, true>());
#pragma unroll
            for (int q = 0; q < PPL; ++q) acc[q] = fma(c6, b6[q], fma(c3, b3[q], acc[q]));
            continue;
        }
        // FP64 grids: every pair in FP64.  The expanded FP64 r^2 of the near
        // pairs of FP32 grids, u = 1/r^2 from the FP32 reciprocal refined in
        // FP64, FP64 sums of u^3 and u^6 scaled by 4 eps sigma^6 / sigma^12
        // per segment.  Far records (inside r_c for every point, farther than
        // kMargin from all) need neither the clamp nor the cutoff test, shell
        // records only the test, near records (an atom within R of the tile
        // centre) both.
        // One Newton step: the FP32 reciprocal of the rounded r^2 is within
        // e0 <= 2^-24 + 2^-23 of 1/r^2, and y + y (1 - r^2 y) leaves 1/r^2
        // (1 - e0^2) plus one FP64 rounding, so u is low by at most 3.2e-14
        // relative and a pair's energy is off by at most 2e-13 of |4 eps
        // sigma^12 u^6| + |4 eps sigma^6 u^3| -- the size of the rounding the
        // expanded r^2 already leaves (|a|^2 + |p|^2 - 2 p.a cancels ~13 FP64
        // ulps of |a|^2).  Near records take two steps.
        const double sig6 = P.sig2 * P.sig2 * P.sig2;
        const double c6 = P.eps4 * sig6 * sig6, c3 = -P.eps4 * sig6;
        double b3[PPL], b6[PPL];
#pragma unroll
        for (int q = 0; q < PPL; ++q) b3[q] = b6[q] = 0.0;
        auto fp64_pair = [&](const double4 &A, auto clamp, auto test) {
            double r2[PPL];
            pair_r2<PPL, XL>(A, mx, my, mz, p2, r2);
#pragma unroll
            for (int q = 0; q < PPL; ++q) {
                const double rr = decltype(clamp)::value ? fmax(r2[q], min_r2) : r2[q];  // :77
                double y = (double)__builtin_amdgcn_rcpf((float)rr);
                y = fma(fma(-rr, y, 1.0), y, y);
                double y3 = y * y * y;
                if (decltype(test)::value) y3 = rr < rc2 ? y3 : 0.0;  // src/computegridsheet.h:79
                b3[q] += y3;
                b6[q] = fma(y3, y3, b6[q]);
            }
        };
        using no = std::integral_constant<bool, false>;
        using yes = std::integral_constant<bool, true>;
        for (int k = fb; k < fe; ++k) fp64_pair(L[k], no(), no());
        // wave lists of an FP64 tile hold as many entries as an FP32 tile's
        // (the far + shell part of the FP32 list), which keeps four
        // workgroups per CU; a segment with more shell records than that
        // takes the tile-level path
        const int wcap = mixed_far_shell_cap(cap);
        if (ws && se - sb <= wcap) {
            // shell records against the wave's own sphere (as for FP32
            // grids): wave-far ones need no cutoff test, those beyond r_c of
            // every point of the wave contribute exactly 0 and are skipped
            int nwf, nws;
            wave_shell_lists(J, Ls, capA, wcap, sb, se, wsph, wl, nwf, nws);
            const char *__restrict__ Lb = reinterpret_cast<const char *>(L);
            for (int e = 0; e < nwf; ++e)
                fp64_pair(*reinterpret_cast<const double4 *>(Lb + wl[e]), no(), no());
            for (int e = 0; e < nws; ++e)
                fp64_pair(*reinterpret_cast<const double4 *>(Lb + wl[wcap - 1 - e]), no(), yes());
        } else {
            for (int k = sb; k < se; ++k) fp64_pair(Ls[-k], no(), yes());
        }
        // Near records (an atom within R of the tile centre, where a grid
        // point may sit on the atom): r^2 from the FP64 differences, as
        // src/computegridsheet.h:71-75 forms it, so that a point on an atom
        // gives exactly 0 and the reference's clamp (the expanded form would
        // leave a rounding residue of either sign there).
        for (int k = nb; k < ne; ++k) {
            const double4 A = Ln[k];
#pragma unroll
            for (int q = 0; q < PPL; ++q) {
                const double dx = px[q] - A.x, dy = py[q] - A.y, dz = pz[q] - A.z;
                const double rr = fmax(dx * dx + dy * dy + dz * dz, min_r2);  // :77
                double y = (double)__builtin_amdgcn_rcpf((float)rr);
                y = fma(fma(-rr, y, 1.0), y, y);
                y = fma(fma(-rr, y, 1.0), y, y);
                double y3 = y * y * y;
                y3 = rr < rc2 ? y3 : 0.0;  // src/computegridsheet.h:79
                b3[q] += y3;
                b6[q] = fma(y3, y3, b6[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < PPL; ++q) acc[q] = fma(c6, b6[q], fma(c3, b3[q], acc[q]));
    }
}

// Energies of one grid tile (tile coordinates tc), candidates streamed from
// src / tbeg (the framework's near list, or a region's sub-list of it).
template <int PPL, int MATH>
__device__ void grid_tile(const JobDev *__restrict__ J, const int tc[3],
                          const double4 *__restrict__ src, const uint32_t *__restrict__ sidx,
                          const int *__restrict__ tbeg, int cap,
                          double4 *__restrict__ s_list, int *__restrict__ s_seg,
                          int *__restrict__ s_cnt, uint16_t *__restrict__ s_wl) {
    const int tid = threadIdx.x;
    int i0[3], i1[3];
    double lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        i0[d] = tc[d] * J->tile[d];
        const int e = i0[d] + J->tile[d];
        i1[d] = (e < J->n[d] ? e : J->n[d]) - 1;
        // src/writegrid.cu:122-127
        lo[d] = 1.0 * i0[d] / (J->n[d] - 1);
        hi[d] = 1.0 * i1[d] / (J->n[d] - 1);
    }
    const Tile tl = tile_sphere(J->t, lo, hi);
    double px[PPL], py[PPL], pz[PPL], acc[PPL];
    long long oidx[PPL];
    bool ok[PPL];
    const int t0 = J->tile[0], t1 = J->tile[1], t2 = J->tile[2];
    // Quadrant layout (FP32 and FP64 grids, even t1 and t2, t0 divisible by PPL): wave w
    // owns the (j, k) quadrant (w & 1, w >> 1) of the tile and all of its i;
    // lane l holds points i = il + q * t0 / PPL.  Otherwise point p = tid +
    // 256 q of the tile in k-fastest order.
    const int wave = tid >> 6, lane = tid & 63;
    // (the wave lists hold 16-bit LDS byte offsets into the far/shell part)
    const bool quad = (MATH == MATH_MIXED || MATH == MATH_FP64) && t1 % 2 == 0 &&
                      t2 % 2 == 0 && t0 % PPL == 0 &&
                      near_region_start<MATH>(cap) * (int)sizeof(double4) <= 65536;
    const int h1 = t1 / 2, h2 = t2 / 2;
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        int gi, gj, gk;
        if (quad) {
            gi = i0[0] + lane / (h1 * h2) + q * (t0 / PPL);
            gj = i0[1] + (wave & 1) * h1 + (lane / h2) % h1;
            gk = i0[2] + (wave >> 1) * h2 + lane % h2;
        } else {
            const int p = q * kWG + tid;
            gi = i0[0] + p / (t1 * t2);
            gj = i0[1] + (p / t2) % t1;
            gk = i0[2] + p % t2;
        }
        ok[q] = gi <= i1[0] && gj <= i1[1] && gk <= i1[2];
        if (!ok[q]) {
            gi = i0[0];
            gj = i0[1];
            gk = i0[2];
        }
        frac_to_cart(J->t, 1.0 * gi / (J->n[0] - 1), 1.0 * gj / (J->n[1] - 1),
                     1.0 * gk / (J->n[2] - 1), px[q], py[q], pz[q]);
        px[q] -= tl.cx;
        py[q] -= tl.cy;
        pz[q] -= tl.cz;
        oidx[q] = (long long)gk + (long long)gj * J->n[2] + (long long)gi * J->n[1] * J->n[2];
        acc[q] = 0.0;
    }
    // A lane's points differ only in i (quadrant layout, or p = tid + 256 q
    // with t1 * t2 dividing 256), so on a cell whose first vector lies along
    // x (t_yx = t_zx = 0, as cell parameters give it) only in x.
    const bool xl = PPL > 1 && (quad || kWG % (t1 * t2) == 0) && J->t[3] == 0.0 &&
                    J->t[6] == 0.0;
    // The sphere around this wave's valid points (quadrant layout): its
    // index box is the tile's i range times the quadrant's j and k ranges.
    bool ws = false, idle = false;
    double4 wsph = make_double4(0.0, 0.0, 0.0, 0.0);
    if (quad) {
        const int jb = i0[1] + (wave & 1) * h1, kb = i0[2] + (wave >> 1) * h2;
        const int je = jb + h1 - 1 < i1[1] ? jb + h1 - 1 : i1[1];
        const int ke = kb + h2 - 1 < i1[2] ? kb + h2 - 1 : i1[2];
        idle = jb > i1[1] || kb > i1[2];  // no point of this wave is on the grid
        if (!idle) {
            const double wlo[3] = {lo[0], 1.0 * jb / (J->n[1] - 1), 1.0 * kb / (J->n[2] - 1)};
            const double whi[3] = {hi[0], 1.0 * je / (J->n[1] - 1), 1.0 * ke / (J->n[2] - 1)};
            const Tile w = tile_sphere(J->t, wlo, whi);
            wsph = make_double4(w.cx - tl.cx, w.cy - tl.cy, w.cz - tl.cz, w.R);
            ws = true;
        }
    }
    int cur_t = 0, cur_pos = uniform(tbeg[0]);
    do {
        const int nseg =
            fill_list<MATH>(J, src, sidx, tbeg, tl, cur_t, cur_pos, cap, s_list, s_seg, s_cnt);
        __syncthreads();
        if (!idle) {
            if (xl)
                eval_list<PPL, MATH, true>(J, s_list, cap, s_seg, nseg, px, py, pz, acc, tl.R, ws,
                                           wsph, s_wl);
            else
                eval_list<PPL, MATH, false>(J, s_list, cap, s_seg, nseg, px, py, pz, acc, tl.R,
                                            ws, wsph, s_wl);
        }
        __syncthreads();
    } while (cur_t < J->ntypes);
#pragma unroll
    for (int q = 0; q < PPL; ++q) {
        if (!ok[q]) continue;
        if (MATH == MATH_COUNT) {
            ((uint32_t *)J->out)[oidx[q]] = (uint32_t)acc[q];
        } else if (MATH == MATH_OVERLAP) {
            ((double *)J->out)[oidx[q]] = acc[q] > 0.0 ? 0.0 : 1.0;  // src/computegridsheet.h:159-162
        } else if (J->mode == OUT_F64) {
            ((double *)J->out)[oidx[q]] = acc[q];
        } else {
            const double v = acc[q] < (double)FLT_MAX ? acc[q] : (double)FLT_MAX;
            ((float *)J->out)[oidx[q]] = (float)v;
        }
    }
}

// Sub-list of the near list for a region of tiles: every image atom within
// r_c + R of the region centre c, where R bounds |c_t - c| + R_t over the
// region's tiles.  Any atom a tile's own fill would keep (d_t - R_t < r_c)
// therefore lies in the sub-list, and the sub-list keeps the near-list order,
// so every tile builds exactly the list it would build from the near list.
// The sub-list holds near-list indices (4 bytes a record): the records
// themselves are read from the near list, which every workgroup on a
// framework shares (it stays in L2), while each workgroup's sub-list is its
// own.  Returns false (and builds nothing usable) if the sub-list exceeds
// cap_s.
__device__ bool build_region_list(const JobDev *__restrict__ J, double cx, double cy, double cz,
                                  double in2, uint32_t *__restrict__ out, int cap_s,
                                  int *__restrict__ rbeg, int *__restrict__ s_cnt) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int n = 0;
    for (int t = 0; t < J->ntypes; ++t) {
        if (tid == 0) rbeg[t] = n;
        const int b = uniform(J->type_begin[t]), e = uniform(J->type_begin[t + 1]);
        for (int pos = b; pos < e; pos += kWG) {
            const int idx = pos + tid;
            bool keep = false;
            double4 A = make_double4(0.0, 0.0, 0.0, 0.0);
            if (idx < e) {
                A = J->near[idx];
                const double dx = A.x - cx, dy = A.y - cy, dz = A.z - cz;
                keep = dx * dx + dy * dy + dz * dz < in2;
            }
            const uint64_t bk = __ballot(keep);
            if (lane == 0) s_cnt[wave] = __popcll(bk);
            __syncthreads();
            int off = n, tot = 0;
#pragma unroll
            for (int w = 0; w < kWG / 64; ++w) {
                if (w < wave) off += s_cnt[w];
                tot += s_cnt[w];
            }
            __syncthreads();
            if (n + tot > cap_s) return false;  // uniform
            if (keep) out[off + lanes_below(bk)] = (uint32_t)idx;
            n = uniform(n + tot);
        }
    }
    if (tid == 0) rbeg[J->ntypes] = n;
    // the sub-list is read back by every wave of the workgroup
    __threadfence();
    __syncthreads();
    return true;
}

template <int PPL, int MATH>
__global__ __launch_bounds__(kWG) void k_tiles(const JobDev *__restrict__ jobs, int njobs,
                                               long long total_items,
                                               unsigned long long *__restrict__ counter,
                                               int cap, uint32_t *__restrict__ scratch,
                                               int scratch_cap) {
    extern __shared__ double4 s_list[];
    __shared__ int s_seg[kSegInts * kMaxSeg];
    __shared__ int s_cnt[7 * (kWG / 64)];  // fill: 3 counts + 4 stop fields per wave
    __shared__ int s_rbeg[kMaxTypes + 1];
    __shared__ double s_red[kWG / 64];
    __shared__ long long s_item;
    const int tid = threadIdx.x;
    uint32_t *__restrict__ slice =
        scratch ? scratch + (size_t)blockIdx.x * (size_t)scratch_cap : nullptr;
    // per-wave shell offset lists after the LDS list (FP32 and FP64 grids)
    uint16_t *__restrict__ s_wl = (MATH == MATH_MIXED || MATH == MATH_FP64)
        ? reinterpret_cast<uint16_t *>(s_list + cap) + (tid >> 6) * mixed_far_shell_cap(cap)
        : nullptr;

    for (;;) {
        if (tid == 0) s_item = (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        const long long item = s_item;
        __syncthreads();
        if (item >= total_items) break;
        int jl = 0, jh = njobs - 1;
        while (jl < jh) {
            const int mid = (jl + jh + 1) >> 1;
            if (jobs[mid].item_begin <= item) jl = mid;
            else jh = mid - 1;
        }
        const JobDev *__restrict__ J = jobs + jl;
        const long long local = item - J->item_begin + J->local0;

        // ---- one region of cells (grid tiles or Widom bins); index
        //      k + j * nz + i * ny * nz over the job's regions ----
        const bool grid = J->kind == JOB_GRID;
        int ncell[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) ncell[d] = grid ? J->ntiles[d] : J->nbins[d];
        long long rem = local;
        int rg[3], t0[3], tn[3];
        rg[2] = (int)(rem % J->nregions[2]);
        rem /= J->nregions[2];
        rg[1] = (int)(rem % J->nregions[1]);
        rg[0] = (int)(rem / J->nregions[1]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            t0[d] = rg[d] * J->region[d];
            const int e = t0[d] + J->region[d];
            tn[d] = (e < ncell[d] ? e : ncell[d]) - t0[d];
        }
        const int nt = tn[0] * tn[1] * tn[2];
        // fractional box of cell tc: a tile's grid points, or a bin
        auto cell_box = [&](const int tc[3], double lo[3], double hi[3]) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (grid) {
                    const int a = tc[d] * J->tile[d];  // src/writegrid.cu:122-127
                    const int b = a + J->tile[d];
                    lo[d] = 1.0 * a / (J->n[d] - 1);
                    hi[d] = 1.0 * ((b < J->n[d] ? b : J->n[d]) - 1) / (J->n[d] - 1);
                } else {
                    lo[d] = (double)tc[d] / J->nbins[d];
                    hi[d] = (double)(tc[d] + 1) / J->nbins[d];
                }
            }
        };
        auto cell_of = [&](int ti, int tc[3]) {
            tc[0] = t0[0] + ti / (tn[1] * tn[2]);
            tc[1] = t0[1] + (ti / tn[2]) % tn[1];
            tc[2] = t0[2] + ti % tn[2];
        };
        const double4 *src = J->near;
        const uint32_t *sidx = nullptr;
        const int *tbeg = J->type_begin;
        if (nt > 1 && slice && J->ntypes <= kMaxTypes) {
            // region centre: middle of the region's box
            double lo[3], hi[3], clo[3], chi[3];
            int tc[3];
            cell_of(0, tc);
            cell_box(tc, lo, clo);
            cell_of(nt - 1, tc);
            cell_box(tc, chi, hi);
            double cx, cy, cz;
            frac_to_cart(J->t, 0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]),
                         0.5 * (lo[2] + hi[2]), cx, cy, cz);
            // R = max over cells of |c_t - c| + R_t (one cell per thread)
            double r = 0.0;
            for (int ti = tid; ti < nt; ti += kWG) {
                cell_of(ti, tc);
                double tlo[3], thi[3];
                cell_box(tc, tlo, thi);
                const Tile ts = tile_sphere(J->t, tlo, thi);
                const double dx = ts.cx - cx, dy = ts.cy - cy, dz = ts.cz - cz;
                const double rt = sqrt(dx * dx + dy * dy + dz * dz) + ts.R;
                r = rt > r ? rt : r;
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const double v = __shfl_xor(r, o, 64);
                r = v > r ? v : r;
            }
            if ((tid & 63) == 0) s_red[tid >> 6] = r;
            __syncthreads();
            double R = s_red[0];
#pragma unroll
            for (int w = 1; w < kWG / 64; ++w) R = s_red[w] > R ? s_red[w] : R;
            R = R * (1.0 + 1e-12) + kMargin;
            __syncthreads();
            const double rr = J->rc + R + kMargin;
            if (build_region_list(J, cx, cy, cz, rr * rr, slice, scratch_cap, s_rbeg, s_cnt)) {
                sidx = slice;
                tbeg = s_rbeg;
            }
        }
        for (int ti = 0; ti < nt; ++ti) {
            int tc[3];
            cell_of(ti, tc);
            if (grid) {
                grid_tile<PPL, MATH>(J, tc, src, sidx, tbeg, cap, s_list, s_seg, s_cnt, s_wl);
                continue;
            }
            // ---- one Widom bin: every insertion whose point falls in it ----
            // (at most 2 per lane: bins hold a few hundred insertions).  A
            // chunk's slots are c + q*256 + tid; a wave none of whose lanes
            // holds a second slot evaluates one point per lane, and a wave
            // with no slot at all skips the chunk (the energies do not depend
            // on the points per lane).  Each pass adds its near records, then
            // its far and shell records, to the insertion's energy.
            //
            // Saturation (J->sat_e > 0, jobs whose energies only feed the
            // Boltzmann sums): no pair energy is below -eps_ij, and none of a
            // far or shell record (beyond r_FP64 of every point) below
            // -fs_floor eps_ij, so once an insertion's partial energy plus
            // those floors for every record not yet evaluated (-eps_ij of its
            // type for records not yet classified) reaches sat_e = 800 T, its exact
            // energy is at least 800 T as well, exp(-E/T) is exactly 0 in
            // FP64 for both, and neither sum changes.  A wave whose valid insertions are all
            // saturated stops evaluating them; the stored partial energy
            // (>= 800 T) then stands in for the exact one.
            constexpr int PW = PPL > 2 ? 2 : PPL;
            // One pass of this wave's P points (px.., valid where ok) over the
            // LDS list: the near records, then, unless every valid point is
            // saturated by then, the far and shell records.  out = prev + the
            // pass's terms (the terms alone in the first pass).  Returns false,
            // out untouched, if the points were saturated before the pass:
            // their stored energies stand.
            auto one_pass = [&](auto pw, const auto &px, const auto &py, const auto &pz,
                                const auto &ok, const auto &prev, auto &out, const Tile &tl,
                                int nseg, bool first, double lb_near, double lb_fs,
                                double lb_later) {
                constexpr int P = decltype(pw)::value;
                double acc[P];
#pragma unroll
                for (int q = 0; q < P; ++q) acc[q] = 0.0;
                const bool sat = J->sat_e > 0.0;
                auto saturated = [&](double lb) {
                    bool all = true;
#pragma unroll
                    for (int q = 0; q < P; ++q)
                        if (ok[q]) all = all && prev[q] + acc[q] + lb >= J->sat_e;
                    return __all(all) != 0;
                };
                if (sat && !first && saturated(lb_near + lb_fs + lb_later)) return false;
                eval_list<P, MATH, false>(J, s_list, cap, s_seg, nseg, px, py, pz, acc, tl.R,
                                          false, make_double4(0.0, 0.0, 0.0, 0.0), nullptr, 1);
                if (!(sat && saturated(lb_fs + lb_later)))
                    eval_list<P, MATH, false>(J, s_list, cap, s_seg, nseg, px, py, pz, acc,
                                              tl.R, false, make_double4(0.0, 0.0, 0.0, 0.0),
                                              nullptr, 2);
#pragma unroll
                for (int q = 0; q < P; ++q) out[q] = first ? acc[q] : prev[q] + acc[q];
                return true;
            };
            // A chunk of up to P * 256 of the bin's insertions, slots
            // c + q * 256 + tid: points generated and energies loaded and
            // stored for this pass.
            auto chunk = [&](auto pw, uint64_t c, uint64_t end, const Tile &tl, int nseg,
                             bool first, double lb_near, double lb_fs, double lb_later) {
                constexpr int P = decltype(pw)::value;
                double px[P], py[P], pz[P], prev[P], out[P];
                uint32_t ins[P];
                bool ok[P];
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    const uint64_t p = c + (uint64_t)(q * kWG + tid);
                    ok[q] = p < end;
                    ins[q] = J->sorted[ok[q] ? p : c];
                    double f[3];
                    widom_point(J->seed, J->stream, ins[q], f);
                    frac_to_cart(J->t, f[0], f[1], f[2], px[q], py[q], pz[q]);
                    px[q] -= tl.cx;
                    py[q] -= tl.cy;
                    pz[q] -= tl.cz;
                    prev[q] = first ? 0.0 : J->e_out[ins[q]];
                }
                if (!one_pass(pw, px, py, pz, ok, prev, out, tl, nseg, first, lb_near, lb_fs,
                              lb_later))
                    return;
#pragma unroll
                for (int q = 0; q < P; ++q)
                    if (ok[q]) J->e_out[ins[q]] = out[q];
            };
            // A bin whose insertions fit one chunk (at most PW * 256, most
            // bins) keeps each lane's points and partial energies in
            // registers across the passes and writes the energies once; the
            // passes add the same terms in the same order as the chunked
            // form, which stores and reloads them after every pass.
            auto resident = [&](auto pw, const double (&X)[PW], const double (&Y)[PW],
                                const double (&Z)[PW], const bool (&OK)[PW], double (&T)[PW],
                                const Tile &tl, int nseg, bool first, double lb_near,
                                double lb_fs, double lb_later) {
                constexpr int P = decltype(pw)::value;
                double px[P], py[P], pz[P], prev[P], out[P];
                bool ok[P];
#pragma unroll
                for (int q = 0; q < P; ++q) {
                    px[q] = X[q];
                    py[q] = Y[q];
                    pz[q] = Z[q];
                    ok[q] = OK[q];
                    prev[q] = first ? 0.0 : T[q];
                }
                if (!one_pass(pw, px, py, pz, ok, prev, out, tl, nseg, first, lb_near, lb_fs,
                              lb_later))
                    return;
#pragma unroll
                for (int q = 0; q < P; ++q) T[q] = out[q];
            };
            const int b = (tc[0] * J->nbins[1] + tc[1]) * J->nbins[2] + tc[2];
            const uint32_t s0 = J->bin_start[b], s1 = J->bin_start[b + 1];
            if (s0 == s1) continue;
            double lo[3], hi[3];
            cell_box(tc, lo, hi);
            const Tile tl = tile_sphere(J->t, lo, hi);
            // one chunk: this wave's points per lane (wave-uniform; 0 for a
            // wave with no slot); several chunks: -1.  Only the FP32-grid /
            // Widom pair math keeps bins resident: Widom jobs never run with
            // the other variants, whose registers the resident state would
            // only crowd (FP64 grids would drop to three waves per SIMD).
            const uint64_t wv0 = s0 + (uint64_t)(tid & ~63);
            const int mode = (MATH != MATH_MIXED || s1 - s0 > (uint32_t)(PW * kWG)) ? -1
                             : wv0 >= s1                     ? 0
                             : (PW > 1 && wv0 + kWG >= s1)   ? 1
                                                             : PW;
            double rx[PW], ry[PW], rz[PW], rt[PW];
            uint32_t rins[PW];
            bool rok[PW];
#pragma unroll
            for (int q = 0; q < PW; ++q) {
                rx[q] = ry[q] = rz[q] = rt[q] = 0.0;
                rins[q] = 0;
                rok[q] = false;
                if (q < mode) {
                    const uint64_t p = s0 + (uint64_t)(q * kWG + tid);
                    rok[q] = p < s1;
                    rins[q] = J->sorted[rok[q] ? p : s0];
                    double f[3];
                    widom_point(J->seed, J->stream, rins[q], f);
                    frac_to_cart(J->t, f[0], f[1], f[2], rx[q], ry[q], rz[q]);
                    rx[q] -= tl.cx;
                    ry[q] -= tl.cy;
                    rz[q] -= tl.cz;
                }
            }
            int cur_t = 0, cur_pos = uniform(tbeg[0]);
            bool first = true;
            do {
                const int nseg = fill_list<MATH>(J, src, sidx, tbeg, tl, cur_t, cur_pos, cap, s_list,
                                                 s_seg, s_cnt);
                __syncthreads();
                // -eps_ij for every record of this pass (near; far and shell)
                // and every source record after it
                double lb_near = 0.0, lb_fs = 0.0, lb_later = 0.0;
                if (J->sat_e > 0.0) {
                    for (int sg = 0; sg < nseg; ++sg) {
                        const int *g = s_seg + kSegInts * sg;
                        const double eps = 0.25 * J->types[uniform(g[0])].eps4;
                        lb_fs -= eps * J->fs_floor *
                                 (double)(uniform(g[2]) - uniform(g[1]) + uniform(g[4]) -
                                                uniform(g[3]));
                        lb_near -= eps * (double)(uniform(g[6]) - uniform(g[5]));
                    }
                    // source records after this pass, each at -eps of its type:
                    // the rest of the current type, then every later type
                    if (cur_t < J->ntypes) {
                        lb_later = -0.25 * J->types[cur_t].eps4 *
                                   (double)(uniform(tbeg[cur_t + 1]) - cur_pos);
                        for (int u = cur_t + 1; u < J->ntypes; ++u)
                            lb_later -= 0.25 * J->types[u].eps4 *
                                        (double)(uniform(tbeg[u + 1]) - uniform(tbeg[u]));
                    }
                }
                if (mode == PW)
                    resident(std::integral_constant<int, PW>(), rx, ry, rz, rok, rt, tl, nseg,
                             first, lb_near, lb_fs, lb_later);
                else if (PW > 1 && mode == 1)
                    resident(std::integral_constant<int, 1>(), rx, ry, rz, rok, rt, tl, nseg,
                             first, lb_near, lb_fs, lb_later);
                for (uint64_t c = s0; mode < 0 && c < s1; c += (uint64_t)(PW * kWG)) {
                    const uint64_t w0 = c + (uint64_t)(tid & ~63);  // wave-uniform
                    if (w0 >= s1) continue;
                    if (PW > 1 && w0 + kWG >= s1)
                        chunk(std::integral_constant<int, 1>(), c, s1, tl, nseg, first, lb_near,
                              lb_fs, lb_later);
                    else
                        chunk(std::integral_constant<int, PW>(), c, s1, tl, nseg, first, lb_near,
                              lb_fs, lb_later);
                }
                first = false;
                __syncthreads();
            } while (cur_t < J->ntypes);
#pragma unroll
            for (int q = 0; q < PW; ++q)
                if (q < mode && rok[q]) J->e_out[rins[q]] = rt[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block g of G sums exp(-E/T) (src/doHenry.h:107) and E exp(-E/T) over the
// contiguous insertion range [g n / G, (g+1) n / G) in a fixed order (thread
// stride, then a fixed shuffle tree and the four wave sums in order).
__device__ __forceinline__ void widom_partial_block(const double *e, long long n,
                                                   double temperature, long long g, long long G,
                                                   double2 *partials, double *s_b,
                                                   double *s_eb) {
    const long long lo = g * n / G, hi = (g + 1) * n / G;
    double sb = 0.0, seb = 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += kWG) {
        const double E = e[i];
        const double bz = exp(-E / temperature);
        sb += bz;
        seb += (bz == 0.0) ? 0.0 : E * bz;
    }
    sb = wave_sum(sb);
    seb = wave_sum(seb);
    if ((threadIdx.x & 63) == 0) {
        s_b[threadIdx.x >> 6] = sb;
        s_eb[threadIdx.x >> 6] = seb;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        partials[g] = make_double2((s_b[0] + s_b[1]) + (s_b[2] + s_b[3]),
                                   (s_eb[0] + s_eb[1]) + (s_eb[2] + s_eb[3]));
}

// One block of kWG threads: thread t sums partials t, t + kWG, ... in order,
// then a fixed shuffle tree per wave and the wave sums in order -- the same
// summation order on every run.
__device__ __forceinline__ void widom_final_block(const double2 *partials, int nblk, double *out,
                                                  double *s_b, double *s_eb) {
    double sb = 0.0, seb = 0.0;
    for (int g = threadIdx.x; g < nblk; g += kWG) {
        sb += partials[g].x;
        seb += partials[g].y;
    }
    sb = wave_sum(sb);
    seb = wave_sum(seb);
    if ((threadIdx.x & 63) == 0) {
        s_b[threadIdx.x >> 6] = sb;
        s_eb[threadIdx.x >> 6] = seb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = (s_b[0] + s_b[1]) + (s_b[2] + s_b[3]);
        out[1] = (s_eb[0] + s_eb[1]) + (s_eb[2] + s_eb[3]);
    }
}

__global__ __launch_bounds__(kWG) void k_widom_partials(const double *e, long long n,
                                                        double temperature, double2 *partials) {
    __shared__ double s_b[kWG / 64], s_eb[kWG / 64];
    widom_partial_block(e, n, temperature, blockIdx.x, gridDim.x, partials, s_b, s_eb);
}

__global__ __launch_bounds__(kWG) void k_widom_final(const double2 *partials, int nblk,
                                                     double *out) {
    __shared__ double s_b[kWG / 64], s_eb[kWG / 64];
    widom_final_block(partials, nblk, out, s_b, s_eb);
}

// Every Widom run of a batch in one launch: block b serves run b / kReduceBlocks.
__global__ __launch_bounds__(kWG) void k_widom_partials_many(const WidomReduce *runs) {
    __shared__ double s_b[kWG / 64], s_eb[kWG / 64];
    const WidomReduce &R = runs[blockIdx.x / kReduceBlocks];
    widom_partial_block(R.e, R.n, R.temperature, blockIdx.x % kReduceBlocks, kReduceBlocks,
                        R.partials, s_b, s_eb);
}

__global__ __launch_bounds__(kWG) void k_widom_final_many(const WidomReduce *runs) {
    __shared__ double s_b[kWG / 64], s_eb[kWG / 64];
    const WidomReduce &R = runs[blockIdx.x];
    widom_final_block(R.partials, kReduceBlocks, R.out, s_b, s_eb);
}

__global__ __launch_bounds__(kWG) void k_sum_u32(const uint32_t *v, long long n,
                                                 unsigned long long *out) {
    unsigned long long s = 0;
    for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < n;
         i += (long long)gridDim.x * kWG)
        s += v[i];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

__global__ __launch_bounds__(kWG) void k_sum_f64_counts(const double *v, long long n,
                                                        unsigned long long *out) {
    unsigned long long s = 0;
    for (long long i = (long long)blockIdx.x * kWG + threadIdx.x; i < n;
         i += (long long)gridDim.x * kWG)
        s += (unsigned long long)v[i];
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

int grid_blocks(long long n, int per_block, int cap_blocks) {
    long long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap_blocks) b = cap_blocks;
    return (int)b;
}

template <int PPL, int MATH>
hipError_t launch_tiles_t(const JobDev *jobs, int njobs, long long total, unsigned long long *ctr,
                          int wgs, int cap, uint32_t *scratch, int scratch_cap, hipStream_t s) {
    hipLaunchKernelGGL((k_tiles<PPL, MATH>), dim3(wgs), dim3(kWG), tiles_lds_bytes(cap, MATH), s, jobs,
                       njobs, total, ctr, cap, scratch, scratch_cap);
    return hipGetLastError();
}

template <int PPL, int MATH>
int occupancy_t(int cap) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tiles<PPL, MATH>, kWG,
                                                     tiles_lds_bytes(cap, MATH)) != hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_near_count(const NearPlan &p, uint32_t *block_count, hipStream_t s) {
    if (p.nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_near_count, dim3(p.nblocks), dim3(kWG), 0, s, p, block_count);
    return hipGetLastError();
}

hipError_t launch_near_offsets(const NearPlan &p, const uint32_t *block_count,
                               uint32_t *block_off, int *type_begin, hipStream_t s) {
    hipLaunchKernelGGL(k_near_offsets, dim3(1), dim3(kWG), 0, s, p, block_count, block_off,
                       type_begin);
    return hipGetLastError();
}

hipError_t launch_near_write(const NearPlan &p, const uint32_t *block_off, double4 *near,
                             hipStream_t s) {
    if (p.nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_near_write, dim3(p.nblocks), dim3(kWG), 0, s, p, block_off, near);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, long long n,
                           hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_blocks(n, kWG, 4096)), dim3(kWG), 0, s, p, v, n);
    return hipGetLastError();
}

int bin_blocks(uint64_t n, int chunk) { return (int)((n + (uint64_t)chunk - 1) / (uint64_t)chunk); }

hipError_t launch_bin_all(const BinJob *jobs, int njobs, long long total_blocks, int chunk,
                          long long max_bins, hipStream_t s) {
    if (njobs <= 0 || total_blocks <= 0) return hipSuccess;
    if (chunk < kWG || total_blocks > (1LL << 31) - 1 || max_bins <= 0) return hipErrorInvalidValue;
    // histogram in LDS for every job when the largest fits, sized for it
    int lds_bins = max_bins <= kLdsBinsMax ? (int)max_bins : 0;
    if (lds_bins > 16384) {
        // more than 64 KiB of dynamic LDS must be allowed per kernel
        static const bool big_lds_ok =
            hipFuncSetAttribute((const void *)k_bin_count,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                kLdsBinsMax * (int)sizeof(uint32_t)) == hipSuccess &&
            hipFuncSetAttribute((const void *)k_bin_scatter,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                kLdsBinsMax * (int)sizeof(uint32_t)) == hipSuccess;
        if (!big_lds_ok) {
            (void)hipGetLastError();
            lds_bins = 0;
        }
    }
    const size_t lds_bytes = sizeof(uint32_t) * (size_t)(lds_bins > 0 ? lds_bins : 1);
    hipLaunchKernelGGL(k_bin_zero, dim3(njobs), dim3(kWG), 0, s, jobs);
    hipLaunchKernelGGL(k_bin_count, dim3((unsigned)total_blocks), dim3(kWG), lds_bytes, s, jobs,
                       njobs, chunk, lds_bins);
    hipLaunchKernelGGL(k_bin_scan, dim3(njobs), dim3(kWG), 0, s, jobs);
    hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)total_blocks), dim3(kWG), lds_bytes, s, jobs,
                       njobs, chunk, lds_bins);
    return hipGetLastError();
}

size_t tiles_lds_bytes(int cap, int math) {
    size_t b = (size_t)cap * sizeof(double4);
    if (math == MATH_MIXED || math == MATH_FP64)
        b += (size_t)(kWG / 64) * mixed_far_shell_cap(cap) * sizeof(uint16_t);
    return (b + 15) & ~(size_t)15;
}

hipError_t launch_tiles(int math, int ppl, const JobDev *jobs, int njobs, long long total,
                        unsigned long long *ctr, int wgs, int cap, uint32_t *scratch,
                        int scratch_cap, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    if (cap < 2 * kWG || wgs <= 0) return hipErrorInvalidValue;
    if (scratch && scratch_cap < kWG) return hipErrorInvalidValue;
#define ER_TILES(P, M)                                                                     \
    if (ppl == P && math == M)                                                             \
        return launch_tiles_t<P, M>(jobs, njobs, total, ctr, wgs, cap, scratch, scratch_cap, s);
    ER_TILES(1, MATH_MIXED) ER_TILES(2, MATH_MIXED) ER_TILES(4, MATH_MIXED)
    ER_TILES(1, MATH_FP64) ER_TILES(2, MATH_FP64) ER_TILES(4, MATH_FP64)
    ER_TILES(1, MATH_COUNT) ER_TILES(2, MATH_COUNT) ER_TILES(4, MATH_COUNT)
    ER_TILES(1, MATH_OVERLAP) ER_TILES(2, MATH_OVERLAP) ER_TILES(4, MATH_OVERLAP)
#undef ER_TILES
    return hipErrorInvalidValue;
}

int tiles_occupancy(int math, int ppl, int cap) {
#define ER_OCC(P, M) \
    if (ppl == P && math == M) return occupancy_t<P, M>(cap);
    ER_OCC(1, MATH_MIXED) ER_OCC(2, MATH_MIXED) ER_OCC(4, MATH_MIXED)
    ER_OCC(1, MATH_FP64) ER_OCC(2, MATH_FP64) ER_OCC(4, MATH_FP64)
    ER_OCC(1, MATH_COUNT) ER_OCC(2, MATH_COUNT) ER_OCC(4, MATH_COUNT)
    ER_OCC(1, MATH_OVERLAP) ER_OCC(2, MATH_OVERLAP) ER_OCC(4, MATH_OVERLAP)
#undef ER_OCC
    return 1;
}

hipError_t launch_widom_partials(const double *e, long long n, double temperature,
                                 double2 *partials, hipStream_t s) {
    hipLaunchKernelGGL(k_widom_partials, dim3(kReduceBlocks), dim3(kWG), 0, s, e, n, temperature,
                       partials);
    return hipGetLastError();
}

hipError_t launch_widom_reduce_many(const WidomReduce *runs, int nruns, hipStream_t s) {
    if (nruns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_widom_partials_many, dim3((unsigned)nruns * kReduceBlocks), dim3(kWG), 0,
                       s, runs);
    hipLaunchKernelGGL(k_widom_final_many, dim3((unsigned)nruns), dim3(kWG), 0, s, runs);
    return hipGetLastError();
}

hipError_t launch_widom_final(const double2 *partials, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_widom_final, dim3(1), dim3(kWG), 0, s, partials, kReduceBlocks, out);
    return hipGetLastError();
}

hipError_t launch_widom_final_n(const double2 *partials, int nblk, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_widom_final, dim3(1), dim3(kWG), 0, s, partials, nblk, out);
    return hipGetLastError();
}

hipError_t launch_sum_u32(const uint32_t *v, long long n, unsigned long long *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_u32, dim3(grid_blocks(n, kWG, 2048)), dim3(kWG), 0, s, v, n, out);
    return hipGetLastError();
}

hipError_t launch_sum_f64_counts(const double *v, long long n, unsigned long long *out,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_f64_counts, dim3(grid_blocks(n, kWG, 2048)), dim3(kWG), 0, s, v, n,
                       out);
    return hipGetLastError();
}

}  // namespace er
