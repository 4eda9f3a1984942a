// er_runtime.h -- host runtime behind the C ABI: device context with a
// caching HBM pool and admission accounting, resident frameworks, plain
// (one framework per launch) grid and Widom runs, and batched runs captured
// in a hipGraph.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "er_common.h"
#include "er_kernels.h"

namespace er {

// Status codes (mirrored by include/emptyroom.h).
enum Status : int {
    ST_OK = 0,
    ST_ARG = 1,
    ST_IO = 2,
    ST_PARSE = 3,
    ST_MISSING = 4,
    ST_HIP = 5,
    ST_NOMEM = 6,
    ST_STATE = 7,
    ST_CAPACITY = 8,
    ST_NODEVICE = 9,
};

void set_error(const std::string &msg);
const char *last_error();

// Reference FrameworkParticle layout (src/datatypes.h:20-28).
struct Particle {
    double x_f, y_f, z_f, eps, sig2;
};

// Kernel shape knobs.  The list capacity and the tile shape decide the order
// of every FP64 sum, so a framework always runs with the same knobs whether it
// runs alone or inside a batch.
struct Tuning {
    int ppl = 2;                 // points per lane (1, 2 or 4)
    int cap = 1024;              // LDS list capacity (records of 32 B)
    int tile[3] = {8, 8, 8};     // grid tile in points, product = 256 * ppl
    int widom_per_bin = 480;     // largest mean insertions per Widom bin
    double s2_switch = 0.444;    // (sigma/r)^2 above which a pair takes the FP64 path (r < 1.5 sigma)
    double region_len = 4.0;     // edge of a region of grid tiles in Angstrom (0: no regions)
    // grid tiles by size class (spacing thresholds in Angstrom; a zero row
    // means `tile`)
    double class_spacing[2] = {0.15, 0.35};
    int class_tile[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

// Size class (0..2) of a grid of n points on cell t: by the geometric mean
// of the lattice-direction point spacings.
int grid_size_class(const double t[9], const int n[3], const Tuning &tu);
// Tile of such a grid: its class's row, or the default tile.
void grid_tile_shape(const double t[9], const int n[3], const Tuning &tu, int tile[3]);
// ST_OK, or ST_ARG with the error set, for knobs the kernels cannot run with.
int check_tuning(const Tuning &tu);

class Context;

class DevicePool {
  public:
    explicit DevicePool(Context *ctx) : ctx_(ctx) {}
    ~DevicePool();
    // Returns nullptr and sets the error if the cap would be exceeded.
    void *alloc(size_t bytes);
    void free(void *p);
    size_t in_use() const { return in_use_; }
    size_t reserved() const { return reserved_; }
    void trim();

  private:
    static size_t round_size(size_t b);
    Context *ctx_;
    std::map<size_t, std::vector<void *>> cache_;
    std::map<void *, size_t> live_;
    size_t in_use_ = 0, reserved_ = 0;
    std::mutex mu_;
};

class Context {
  public:
    int device = 0;
    hipStream_t stream = nullptr;
    size_t hbm_cap = 0;  // 0 = no cap beyond the device
    int n_cu = 0;
    Tuning tuning;
    DevicePool pool{this};
    unsigned long long *d_counter = nullptr;  // work counter of plain launches
    double2 *d_partials = nullptr;
    double *d_sums = nullptr;
    uint32_t *d_scratch = nullptr;  // region sub-lists, scratch_wgs * kRegionRecords indices
    int scratch_wgs = 0;
    void *h_stage = nullptr;       // pinned host staging of pocket labelling results
    size_t h_stage_bytes = 0;
    ~Context();
    // Pinned host buffer of at least `bytes` (kept, grown on demand); nullptr
    // with the error set if the allocation fails.
    void *host_stage(size_t bytes);
};

int context_create(int device, size_t hbm_cap, Context **out);
// Persistent workgroups a launch of the tile kernel may use for `items` work
// items, given the tuning's occupancy and a scratch of `scratch_wgs` slices.
int tiles_workgroups(const Context *ctx, int math, long long items, int scratch_wgs);
// Region scratch for every workgroup the current tuning can keep resident
// (n_cu * occupancy); grows *buf if needed.  Called at context creation and
// after a tuning change; never during graph capture.
int scratch_ensure(Context *ctx, hipStream_t s, uint32_t **buf, int *wgs);

struct FrameworkDev {
    Context *ctx = nullptr;
    int n_atoms = 0, ntypes = 0;
    double t[9], inv_t[9], rc2 = 0, rc = 0, volume = 0;
    double s2_switch = 0;  // (sigma / r_FP64)^2 the types' switch radii were built with
    int rep[3] = {0, 0, 0};
    long long n_images = 0;
    size_t near_capacity = 0;
    std::vector<TypeParams> types;
    AtomDev *d_atoms = nullptr;
    int *d_type_atom0 = nullptr, *d_type_block0 = nullptr, *d_type_begin = nullptr;
    TypeParams *d_types = nullptr;
    double4 *d_near = nullptr;
    uint32_t *d_block_count = nullptr, *d_block_off = nullptr;
    std::vector<void *> owned;
    ~FrameworkDev();
};

int framework_upload(Context *ctx, const Particle *atoms, int n, const double t[9],
                     const int rep[3], double rc2, FrameworkDev **out);

struct GridDev {
    Context *ctx = nullptr;
    const FrameworkDev *fw = nullptr;
    int n[3] = {0, 0, 0};
    int precision = 0;  // 0 FP32, 1 FP64
    void *d_data = nullptr;
    size_t bytes() const;
    ~GridDev();
};

struct WidomResult {
    double sum_b = 0, sum_eb = 0;
    uint64_t n = 0;
};

// Plain path: one framework per launch, synchronous.
int grid_compute(Context *ctx, FrameworkDev *fw, const int n[3], int precision, void *host_out,
                 GridDev **resident);
int widom_run(Context *ctx, FrameworkDev *fw, uint64_t n_ins, double temperature, uint64_t seed,
              WidomResult *out, double *energies_host);
// In-cutoff (point, atom, image) triples of a grid / a Widom run.
// per_point_host (nullable) receives each point's count, grid layout.
int count_incut_grid(Context *ctx, FrameworkDev *fw, const int n[3], unsigned long long *out,
                     uint32_t *per_point_host = nullptr);
int count_incut_widom(Context *ctx, FrameworkDev *fw, uint64_t n_ins, uint64_t seed,
                      unsigned long long *out);
// The grid of grid_compute written into caller-owned device memory d_out
// (n[0] n[1] n[2] floats or doubles), with no device allocation: the work
// counter, the job descriptor and the region sub-lists live in a fixed
// device array of the library (er_noalloc.hip).  Same values as
// grid_compute bit for bit.  Calls that overlap on one device race.
int grid_compute_noalloc(Context *ctx, FrameworkDev *fw, const int n[3], int precision,
                         void *d_out);
int check_grid_dims(const int n[3]);
// Hard-sphere accessibility grid (src/computegridsheet.h:101-165): 1 where no
// atom image with offsets in [-1, 1]^3 lies closer than sigma_ij, else 0;
// grid points f = i / (N - 1), layout k + j*nz + i*ny*nz.
int accessibility_grid(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                       const int n[3], double *host_out);
// The reference's direct method (src/computegridsheet.h:30-99,
// src/doHenry.h:29-135) on the GPU in FP64, every (point, image, atom)
// triple, atoms in the caller's order (er_direct.hip).  The grid is FP64;
// resident (nullable) keeps it on the device.
int direct_grid(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                const int rep[3], double rc2, const int n[3], double *host_out,
                GridDev **resident);
int direct_widom(Context *ctx, const Particle *atoms, int n_atoms, const double t[9],
                 const int rep[3], double rc2, uint64_t n_ins, double temperature, uint64_t seed,
                 WidomResult *out, double *energies_host);

// Device buffers and job descriptors of one Widom job.
struct WidomBuffers {
    int nbins[3] = {1, 1, 1};
    uint32_t *count = nullptr, *start = nullptr, *cursor = nullptr, *sorted = nullptr;
    double *e = nullptr;
    BinJob job{};               // host copy of the binning job (source of the upload)
    BinJob *d_job = nullptr;    // device copy
};
BinJob make_bin_job(const WidomBuffers &wb, uint64_t n_ins, uint64_t seed, uint32_t stream_id);
void choose_widom_bins(const FrameworkDev *fw, uint64_t n_ins, const Tuning &tu, int nb[3]);
void choose_grid_tile(const FrameworkDev *fw, const int n[3], const Tuning &tu, int tile[3]);
// slice_records: records of one workgroup's region scratch slice (regions
// are sized so their expected sub-list fills at most 3/4 of a slice).
void choose_grid_region(const FrameworkDev *fw, const int n[3], const int tile[3],
                        const Tuning &tu, int region[3], int slice_records = kRegionRecords);
JobDev make_grid_job(const FrameworkDev *fw, const int n[3], int precision, void *d_out,
                     const Tuning &tu, int math, int slice_records = kRegionRecords);
void choose_region(const FrameworkDev *fw, const double cell_frac[3], const int ncell[3],
                   const Tuning &tu, long long min_items, int region[3],
                   int slice_records = kRegionRecords);
// min_items: regions of bins are shrunk until the job has at least this many
// work items (kMinGridItems for a job launched alone).  temperature > 0 lets
// the tile kernel stop an insertion once its energy is certain to reach
// 800 T (for jobs whose energies only feed the Boltzmann sums); 0 computes
// every energy in full.
JobDev make_widom_job(const FrameworkDev *fw, uint64_t n_ins, uint64_t seed,
                      const WidomBuffers &wb, const Tuning &tu, long long min_items,
                      double temperature = 0.0);
// Allocates the Widom buffers of a job (owned pointers appended to `owned`).
int widom_prepare(Context *ctx, FrameworkDev *fw, uint64_t n_ins, WidomBuffers *wb,
                  std::vector<void *> *owned);
// Enqueues point generation, binning and the bin-sorted index list on s.
int widom_enqueue_binning(Context *ctx, WidomBuffers &wb, uint64_t n_ins, uint64_t seed,
                          uint32_t stream_id, hipStream_t s);
// Upper bound on the near-list records framework_upload allocates.
size_t framework_near_capacity(int n_atoms, const double t[9], const int rep[3], double rc2);

// Config 5: FP32 grids uploaded from the host, pocket blocking
// (src/pocketblocking.cc:126-391) and grid-interpolated Widom (src/gcmc.cc:145-177).
int grid_upload(Context *ctx, const float *host, const int n[3], GridDev **out);
// Labels (and with block, blocks) many grids together; seg_host (nullable
// array, nullable entries) receives per-grid segment ids.
int pockets_label_many(Context *ctx, GridDev *const *grids, int ng, double temperature,
                       bool block, int *n_pockets, int32_t *const *seg_host);
int pockets_label(Context *ctx, GridDev *g, double temperature, bool block, int *n_pockets,
                  int32_t *seg_host);
int widom_grid_run(Context *ctx, GridDev *g, uint64_t n_ins, double temperature, uint64_t seed,
                   WidomResult *out, double *energies_host);
// Energies at n host positions (fractional, any range) into host out.
int grid_energies(Context *ctx, GridDev *g, const double *frac_host, uint64_t n,
                  double *out_host);
// Segment classification from face links (3 ints each: from, to, axis 1..3,
// from < 0 = no link); returns the number of pockets, -1 on bad links.
int classify_pockets(int nseg, const int *links, long long nlinks, std::vector<uint8_t> *pocket);

// Batched path: many (framework, grid, Widom) tasks in one persistent launch,
// captured in a hipGraph on first run and replayed.
struct Task {
    FrameworkDev *fw = nullptr;
    int grid_n[3] = {0, 0, 0};  // all zero: no grid
    int precision = 0;
    uint64_t n_ins = 0;         // 0: no Widom
    double temperature = 298.0;
    uint64_t seed = 0;
};

class Batch {
  public:
    Context *ctx = nullptr;
    std::vector<Task> tasks;
    std::vector<GridDev *> grids;      // per task (nullptr if no grid)
    std::vector<WidomBuffers> wbufs;   // per task
    std::vector<JobDev> jobs;          // all jobs, cost-ordered
    JobDev *d_jobs = nullptr;
    std::vector<BinJob> bin_jobs;      // Widom binning of every task with insertions
    BinJob *d_bin_jobs = nullptr;
    long long bin_blocks_total = 0, bin_max = 0;
    long long total_items = 0;
    int math = MATH_MIXED;
    double2 *d_partials = nullptr;     // kReduceBlocks per task
    double *d_sums = nullptr;          // 2 per task
    std::vector<WidomReduce> reduce;   // reduction of every task with insertions
    WidomReduce *d_reduce = nullptr;
    unsigned long long *d_counter = nullptr;
    double *h_sums = nullptr;          // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
    // eager runs only: events around each tiled-kernel launch of the last run
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int timed_launches = 0;
    std::vector<void *> owned;
    ~Batch();
    int record(hipStream_t s);  // enqueue the whole batch on s
};

int batch_create(Context *ctx, const Task *tasks, int n, bool use_graph, Batch **out);
int batch_run(Batch *b);   // asynchronous
int batch_wait(Batch *b);  // waits and fetches the Widom sums
int batch_result(Batch *b, int i, WidomResult *out, GridDev **grid);
int batch_set_timing(Batch *b, bool on);
// Milliseconds of the tiled-kernel launches of the last eager run.
int batch_kernel_ms(Batch *b, double *ms, int *launches);

}  // namespace er
