// er_kernels.h -- host-side launchers of the gfx950 kernels in er_kernels.hip.
// Every launcher validates its sizes on the host before it launches and
// returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "er_common.h"

namespace er {

// Pair-math variants of the tiled energy kernel.
enum Math : int { MATH_MIXED = 0, MATH_FP64 = 1, MATH_COUNT = 2, MATH_OVERLAP = 3 };

// Framework atom in fractional coordinates with its type index (atoms are
// grouped by type before upload).
struct AtomDev {
    double fx, fy, fz;
    int type;
    int pad;
};

// --- near-image list (per framework) -------------------------------------
// Enumeration: for each type t, images (reference order) x atoms of type t,
// padded to whole 256-thread blocks per type.
struct NearPlan {
    int ntypes;
    int rep[3];
    double t[9];
    double ext[3];            // fractional slab half-extension r_c*|row_i(T^-1)|
    const AtomDev *atoms;     // n atoms, grouped by type
    const int *type_atom0;    // ntypes + 1 atom offsets
    const int *type_block0;   // ntypes + 1 block offsets of the padded enumeration
    int nblocks;
    int n_images;
};

hipError_t launch_near_count(const NearPlan &p, uint32_t *block_count, hipStream_t s);
hipError_t launch_near_offsets(const NearPlan &p, const uint32_t *block_count,
                               uint32_t *block_off, int *type_begin, hipStream_t s);
hipError_t launch_near_write(const NearPlan &p, const uint32_t *block_off, double4 *near,
                             hipStream_t s);

// --- fills ------------------------------------------------------------------
hipError_t launch_fill_u64(unsigned long long *p, unsigned long long v, long long n,
                           hipStream_t s);

// --- Widom binning (batched) --------------------------------------------------
// jobs: device array of njobs BinJob whose block0 fields number the blocks of
// `chunk` insertions each; total_blocks = sum over jobs of ceil(n / chunk).
// Order: zero, count, scan, scatter (all on one stream).
hipError_t launch_bin_all(const BinJob *jobs, int njobs, long long total_blocks, int chunk,
                          long long max_bins, hipStream_t s);
int bin_blocks(uint64_t n, int chunk);

// --- tiled energy kernel ------------------------------------------------------
// Runs the work items [0, total_items) of the jobs (device array) with a
// persistent grid of `workgroups` blocks pulling items from *counter (which
// must be zero).  cap = LDS list capacity in records.  scratch (nullable)
// holds workgroups * scratch_cap entries of region sub-lists (near-list
// indices); without it every tile streams the framework's whole near list.
hipError_t launch_tiles(int math, int ppl, const JobDev *jobs_dev, int njobs,
                        long long total_items, unsigned long long *counter, int workgroups,
                        int cap, uint32_t *scratch, int scratch_cap, hipStream_t s);
// LDS bytes the tiled kernel needs for a list capacity.
size_t tiles_lds_bytes(int cap, int math);
// Far + shell part of an FP32-grid LDS list of cap records (the rest holds
// near records): both parts hold at least one chunk of kWG records
// (cap >= 2 kWG), so a fill pass that starts empty always makes progress.
__host__ __device__ inline int mixed_far_shell_cap(int cap) {
    const int near_cap = (cap / 3) & ~63;
    return cap - (near_cap < kWG ? kWG : near_cap);
}
// Resident workgroups per CU for (math, ppl, cap).
int tiles_occupancy(int math, int ppl, int cap);

// --- reductions ---------------------------------------------------------------
// Boltzmann sums of one job: partial[g] for g < nblk over fixed index ranges.
constexpr int kReduceBlocks = 240;
hipError_t launch_widom_partials(const double *e, long long n, double temperature,
                                 double2 *partials, hipStream_t s);
// Fixed-order sum of nblk partials (out[0] = sum_b, out[1] = sum_Eb).
hipError_t launch_widom_final_n(const double2 *partials, int nblk, double *out, hipStream_t s);
// out[0] = sum_b, out[1] = sum_Eb: fixed-order sum of the kReduceBlocks partials.
hipError_t launch_widom_final(const double2 *partials, double *out, hipStream_t s);
// One Widom run's reduction: energies e[0, n), kReduceBlocks partials, two sums.
struct WidomReduce {
    const double *e;
    long long n;
    double temperature;
    double2 *partials;
    double *out;
};
// The partials and final sums of nruns runs (device table) in two launches,
// each run summed exactly as launch_widom_partials + launch_widom_final do.
hipError_t launch_widom_reduce_many(const WidomReduce *runs, int nruns, hipStream_t s);
// Sum of n uint32 counts into *out (uint64, must be zero).
hipError_t launch_sum_u32(const uint32_t *v, long long n, unsigned long long *out, hipStream_t s);
// Sum of n FP64-stored integer counts into *out (uint64, must be zero).
hipError_t launch_sum_f64_counts(const double *v, long long n, unsigned long long *out,
                                 hipStream_t s);

}  // namespace er
