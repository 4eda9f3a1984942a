// er_common.h -- definitions shared by the host runtime and the gfx950
// kernels: the Widom point stream, the per-framework device view and the
// work-item descriptor of the tiled energy kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace er {

// ---------------------------------------------------------------------------
// Widom point stream.  Philox4x32-10 (Salmon et al., SC'11) keyed by the job
// seed; insertion i uses counter (lo(i), hi(i), stream, 0) and maps the first
// three output words to (w + 1) * 2^-32, i.e. fractional coordinates in
// (0, 1] like curand_uniform_double in src/doHenry.h:53-55.  Stream 0 is the
// direct-sum Widom run, stream 1 the grid-interpolated one.
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__host__ __device__ inline void widom_point(uint64_t seed, uint32_t stream, uint64_t i,
                                            double f[3]) {
    uint32_t w[4];
    philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), stream, 0u, (uint32_t)seed,
                  (uint32_t)(seed >> 32), w);
    const double inv2_32 = 2.3283064365386962890625e-10;  // 2^-32, exact
    f[0] = ((double)w[0] + 1.0) * inv2_32;
    f[1] = ((double)w[1] + 1.0) * inv2_32;
    f[2] = ((double)w[2] + 1.0) * inv2_32;
}

// ---------------------------------------------------------------------------
// Device data of one uploaded framework.
// ---------------------------------------------------------------------------

// Pair parameters of one (guest, framework-atom-type) combination after
// Lorentz-Berthelot mixing: sig2 = sigma_ij^2, eps4 = 4 eps_ij.  rclose is the
// distance below which a pair is evaluated on the FP64 path.
struct alignas(16) TypeParams {
    double sig2;
    double eps4;
    double rclose;
    double pad;
};

// Work kinds of the tiled energy kernel.
enum JobKind : int { JOB_GRID = 0, JOB_WIDOM = 1 };
// Output modes: FP32 grid (FP32 far-pair LJ, FP64 close pairs and
// accumulation), FP64 grid (every pair on the FP64 path), in-cutoff count.
enum OutMode : int { OUT_F32 = 0, OUT_F64 = 1, OUT_COUNT = 2 };

// One job of the tiled energy kernel: a grid over the unit cell or the Widom
// insertions of one framework.  Plain-old-data, copied to device memory.
struct JobDev {
    double t[9];        // fractional -> Cartesian, row-major (src/Framework.cc:124-132)
    double rc2;         // r_c^2
    double rc;          // r_c
    double min_r2;      // r^2 clamp: 1e-24 grid, 1e-32 Widom
    const double4 *near;      // image atoms near the cell, grouped by type (x, y, z, 0)
    const int *type_begin;    // ntypes + 1 offsets into near
    const TypeParams *types;  // ntypes
    int ntypes;
    int kind;           // JobKind
    int mode;           // OutMode
    int pad0;
    // grid
    int n[3];           // grid points per axis, endpoints included
    int tile[3];        // tile extent in points
    int ntiles[3];      // tiles per axis
    int region[3];      // tiles per region (work item) per axis
    int nregions[3];    // regions per axis
    void *out;          // float*, double* or uint32_t* by mode; index k + j*nz + i*ny*nz
    // Widom
    int nbins[3];       // insertion bins per fractional axis
    int pad1;
    const uint32_t *sorted;     // insertion indices ordered by bin
    const uint32_t *bin_start;  // nbins_total + 1 offsets into sorted
    uint64_t seed;
    uint32_t stream;
    uint32_t pad2;
    double *e_out;      // FP64 energy per insertion index
    double sat_e;       // > 0: stop at a partial energy this high (800 T; see k_tiles)
    double fs_floor;    // far and shell pairs (beyond r_FP64 of every point) add at least
                        // -fs_floor * eps_ij each (<= 1)
    // point boxes (fractional half-extents per axis): a full grid tile's or a
    // bin's, and a full wave quadrant's of a grid tile; a partial one lies in
    // the full one centred on it
    double cell_box[3];
    double wave_box[3];
    // scheduling
    long long item_begin;  // first global work item of this job
    long long n_items;     // grid tiles or Widom bins of this job in this launch
    long long local0;      // first tile / bin of the job covered by this launch
};

// One Widom binning job: insertion points of (seed, stream) sorted into a
// regular grid of fractional bins.  Plain-old-data; kernels take an array.
struct BinJob {
    uint64_t seed;
    uint64_t n;            // insertions
    uint32_t stream;
    int nb[3];             // bins per fractional axis
    uint32_t *count;       // nbins + 1 counts (zeroed by k_bin_zero)
    uint32_t *start;       // nbins + 1 exclusive prefix sums
    uint32_t *cursor;      // nbins fill cursors
    uint32_t *sorted;      // n insertion indices, grouped by bin
    long long block0;      // first binning block of this job in a launch
};

// Kernel launch geometry.
constexpr int kWG = 256;      // threads per workgroup (4 wave64)
constexpr int kMaxSeg = 64;   // list segments per fill pass
constexpr int kMaxTypes = 64; // pair types a region sub-list can index
constexpr int kLdsBinsMax = 32768; // Widom bins a binning block histograms in LDS (128 KiB)
constexpr int kBinChunk = 16384;  // insertions per binning block
constexpr int kRegionRecords = 8192;  // entries of one workgroup's region sub-list (near-list
                                      // indices, 32 KiB)
constexpr long long kMinGridItems = 2048;  // regions per grid before regions shrink

}  // namespace er
