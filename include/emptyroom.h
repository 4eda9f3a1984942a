/*
 * emptyroom.h -- C ABI of libemptyroom.so, the MI355X (gfx950) implementation
 * of EmptyRoom's two GPU hot paths: the guest-framework Lennard-Jones energy
 * grid (writegrid) and Widom-insertion Henry coefficients (henry).
 *
 * Plain C, POD arguments only, no exceptions across the boundary.  Every call
 * returns an er_status (0 = ER_OK); on failure er_last_error() holds a
 * message for the calling thread.  One context serves one host thread at a
 * time; the server API serialises internally.
 *
 * Reference interfaces replaced (paths relative to the EmptyRoom tree):
 *   ComputeGridSheet + the x-sheet loop   src/computegridsheet.h:30-99,
 *                                         src/writegrid.cu:175-271
 *   doHenry + launch + block reduction    src/doHenry.h:29-135,
 *                                         src/henry.cu:111-169
 *   FrameworkParticle                     src/datatypes.h:20-28
 *   host loaders and .out writers         src/Framework.cc, src/Forcefield.cc,
 *                                         src/readsettings.h,
 *                                         src/load_fast_particle_f_array.h,
 *                                         src/write_settings_to_outputfile.h
 */
#ifndef EMPTYROOM_H_
#define EMPTYROOM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum er_status {
    ER_OK = 0,
    ER_E_ARG = 1,       /* invalid argument */
    ER_E_IO = 2,        /* file missing or unreadable */
    ER_E_PARSE = 3,     /* malformed input file */
    ER_E_MISSING = 4,   /* required key, label or mass not found */
    ER_E_HIP = 5,       /* HIP runtime error */
    ER_E_NOMEM = 6,     /* allocation failed or HBM cap reached */
    ER_E_STATE = 7,     /* call not valid in the object's current state */
    ER_E_CAPACITY = 8,  /* server queue full */
    ER_E_NODEVICE = 9   /* no HIP device */
} er_status;

typedef enum er_precision {
    ER_FP32 = 0, /* FP32 grid: FP32 LJ for pairs beyond 1.5 sigma (tunable), FP64 r^2, close pairs and sums */
    ER_FP64 = 1  /* FP64 grid: FP64 sums, far pairs from an FP32 reciprocal plus one Newton step
                    (bound: 2e-13 of the sum of the terms' magnitudes; measured on the batching
                    mix: within 1.3e-10 of max(1,|E|) of the reference's direct method) */
} er_precision;

typedef struct er_context er_context;
typedef struct er_framework er_framework;
typedef struct er_grid er_grid;
typedef struct er_batch er_batch;
typedef struct er_server er_server;
typedef struct er_host_run er_host_run;

/* One framework atom after mixing with the guest: fractional position and
 * eps_ij (K), sigma_ij^2 (A^2).  Same layout as the reference's
 * FrameworkParticle (src/datatypes.h:20-28). */
typedef struct er_particle {
    double x_f, y_f, z_f;
    double eps_with_adsorbate;
    double sig_with_adsorbate_squared;
} er_particle;

/* Widom sums: sum of exp(-E/T) and of E*exp(-E/T) over n insertions, the two
 * fields of HenryStats the host reduces (src/datatypes.h:162-168,
 * src/henry.cu:157-164). */
typedef struct er_widom_result {
    double sum_b;
    double sum_eb;
    uint64_t n;
} er_widom_result;

/* Shape knobs of the tile kernel (grids and direct Widom).  They decide the
 * order of every FP64 sum, so a context keeps them for all of its work. */
typedef struct er_tuning {
    int points_per_lane;   /* 1, 2 or 4 (Widom insertions: at most 2 per lane) */
    int lds_records;       /* LDS candidate-list capacity, 512..4800 records of 32 B */
    int tile[3];           /* grid tile in points, product = 256 * points_per_lane */
    int widom_per_bin;     /* largest mean insertions per Widom bin */
    double s2_switch;      /* (sigma/r)^2 above which an FP32-grid pair uses FP64 */
    double region_len;     /* edge in Angstrom of the tile regions that share one
                              candidate sub-list (0 = every tile scans the whole
                              near list) */
    /* Grid tiles by size class.  A grid's spacing h is the geometric mean of
     * its three lattice-direction point spacings |a_d| / (N_d - 1); class 0
     * is h < class_spacing[0], class 1 is h < class_spacing[1], class 2 the
     * rest.  A class whose class_tile row is all zero uses `tile`; a non-zero
     * row must multiply to 256 * points_per_lane.  The class depends only on
     * the cell and the point counts, so a grid gets the same tile alone, in a
     * batch or in the server. */
    double class_spacing[2];
    int class_tile[3][3];
} er_tuning;

const char *er_last_error(void);
const char *er_version(void);

/* ---- context ------------------------------------------------------------ */
/* hbm_cap_bytes: 0 = whole device; otherwise device allocations of this
 * context (frameworks, grids, scratch) are refused beyond the cap.
 * Frameworks, grids, batches and servers use their context's memory pool
 * and stream: release them before er_context_destroy. */
int er_context_create(int device, size_t hbm_cap_bytes, er_context **out);
int er_context_destroy(er_context *ctx);
int er_context_get_tuning(const er_context *ctx, er_tuning *out);
int er_context_set_tuning(er_context *ctx, const er_tuning *tuning);
/* Host only (no device): the library's default knobs. */
int er_tuning_defaults(er_tuning *out);
/* Host only (no device): the size class (0..2) and tile that `tuning` gives
 * a grid of n points per axis on cell t_matrix (row-major, Cartesian =
 * t_matrix * fractional).  No reference counterpart: the reference launches
 * 16x16-point blocks per grid sheet (src/writegrid.cu:190-197). */
int er_grid_tile_for(const er_tuning *tuning, const double t_matrix[9], const int n[3],
                     int tile_out[3], int *size_class);
int er_context_memory(const er_context *ctx, size_t *in_use, size_t *reserved);

/* ---- frameworks ----------------------------------------------------------
 * Replaces the atom upload of src/writegrid.cu:186-189 / src/henry.cu:133-135
 * and the by-value GridParameters/HenryParameters fields the kernels read
 * (t_matrix, replication factors, r_cutoff_squared; src/datatypes.h:44-110).
 * t_matrix is row-major fractional->Cartesian (src/Framework.cc:124-132);
 * rep[3] are the image half-ranges of the "once" .uc file; rc2 = r_c^2.
 * The device copy stays resident until er_framework_release. */
int er_framework_upload(er_context *ctx, const er_particle *atoms, int n_atoms,
                        const double t_matrix[9], const int rep[3], double rc2,
                        er_framework **out);
int er_framework_release(er_framework *fw);

/* ---- energy grid (plain path) ---------------------------------------------
 * Replaces the whole x-sheet loop of src/writegrid.cu:219-271 around
 * ComputeGridSheet (src/computegridsheet.h:30-99): the energy (K) at every
 * fractional point (i/(nx-1), j/(ny-1), k/(nz-1)), stored at
 * k + j*nz + i*ny*nz (src/computegridsheet.h:94-97).  host_out (nullable)
 * receives float[nx*ny*nz] for ER_FP32 or double[...] for ER_FP64;
 * resident_out (nullable) keeps the device grid.  FP32 values at or above
 * FLT_MAX are stored as FLT_MAX. */
int er_grid_compute(er_context *ctx, er_framework *fw, const int n[3], int precision,
                    void *host_out, er_grid **resident_out);
/* No counterpart in the reference: er_grid_compute for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The grid (n[0] n[1] n[2] floats for
 * ER_FP32, doubles for ER_FP64, layout k + j*N_z + i*N_y*N_z) is written into
 * caller-owned device memory device_out; the call allocates nothing: the
 * work counter, job descriptor and region sub-lists live in a fixed 32 MiB
 * device array of the library (one per device, held until the process ends;
 * no input size needs more).  Values are er_grid_compute's bit for bit.  The
 * call returns when the grid is written.  Calls that overlap on one device
 * (other threads or contexts) share that array and race. */
int er_grid_compute_noalloc(er_context *ctx, er_framework *fw, const int n[3], int precision,
                            void *device_out);
int er_grid_download(er_grid *grid, void *host_out);
int er_grid_info(const er_grid *grid, int n[3], int *precision);
int er_grid_release(er_grid *grid);

/* ---- Widom insertion (plain path) ------------------------------------------
 * Replaces doHenry and its launch (src/doHenry.h:29-135,
 * src/henry.cu:111-164): n_insertions points from the Philox4x32-10 stream
 * keyed by `seed` (the reference draws from cuRAND MTGP32 seeded 1234),
 * guest-framework energy with the r^2 clamp 1e-32, FP64 exp(-E/T) and FP64
 * sums in a fixed order.  energies_out (nullable) receives E (K) per
 * insertion index, every energy computed in full.  Without it, an insertion
 * whose energy is certain to reach 800 T (no pair energy is below -eps_ij)
 * stops early: its exp(-E/T) is exactly 0 either way, so the sums are
 * bit-identical.  K_H and <E> follow from er_henry_result. */
int er_widom(er_context *ctx, er_framework *fw, uint64_t n_insertions, double temperature,
             uint64_t seed, er_widom_result *out, double *energies_out);

/* Fractional coordinates in (0,1] of insertion i of the stream (seed, stream):
 * stream 0 is er_widom's.  Lets a caller replay the exact points of a run
 * (the reference's MTGP32 draws, src/doHenry.h:53-55, cannot be replayed). */
void er_widom_point(uint64_t seed, uint32_t stream, uint64_t i, double f[3]);

/* Number of (point, atom, image) triples with r^2 < r_c^2: the work unit of
 * the roofline figures.  Uses the same tile lists as the energy kernels. */
int er_count_incut_grid(er_context *ctx, er_framework *fw, const int n[3], uint64_t *out);
/* The same count per grid point (uint32, grid layout k + j*N_z + i*N_y*N_z)
 * and its total (nullable): the in-cutoff test of src/computegridsheet.h:79
 * counted point by point. */
int er_count_incut_grid_map(er_context *ctx, er_framework *fw, const int n[3], uint64_t *total,
                            uint32_t *counts_out);
int er_count_incut_widom(er_context *ctx, er_framework *fw, uint64_t n_insertions, uint64_t seed,
                         uint64_t *out);

/* K_H (mol/(kg Pa) and mol/(m^3 Pa)) and <E> (K, kJ/mol) from Widom sums,
 * src/henry.cu:166-174. */
void er_henry_result(double sum_b, double sum_eb, double n_total, double temperature,
                     double density, double *kh_mol_per_kg_pa, double *kh_mol_per_m3_pa,
                     double *e_avg_k, double *e_avg_kj_per_mol);

/* ---- batched path ----------------------------------------------------------
 * No reference counterpart (the reference runs one framework per process).
 * Many (framework, grid, Widom) tasks run as one persistent launch; with
 * use_graph the whole batch is captured in a hipGraph on the first run and
 * replayed afterwards.  Results are identical to the plain path. */
typedef struct er_task {
    er_framework *fw;
    int grid_n[3];          /* all zero: no grid */
    int precision;          /* er_precision of the grid */
    uint64_t n_insertions;  /* 0: no Widom */
    double temperature;
    uint64_t seed;
} er_task;

int er_batch_create(er_context *ctx, const er_task *tasks, int n_tasks, int use_graph,
                    er_batch **out);
int er_batch_run(er_batch *batch);   /* asynchronous */
int er_batch_wait(er_batch *batch);  /* blocks, then results are valid */
/* grid_out (nullable) receives a view owned by the batch (do not release). */
int er_batch_result(er_batch *batch, int task, er_widom_result *out, er_grid **grid_out);
int er_batch_release(er_batch *batch);
/* Eager runs (use_graph = 0) only: bracket every tiled-kernel launch with HIP
 * events on the library's stream; er_batch_kernel_ms then returns the summed
 * milliseconds of those launches in the last run and how many there were. */
int er_batch_set_timing(er_batch *batch, int on);
int er_batch_kernel_ms(er_batch *batch, double *ms, int *launches);

/* ---- pocket blocking and grid-interpolated Widom ---------------------------
 * An FP32 grid (from er_grid_compute, a batch, or er_grid_upload) is
 * labelled as src/pocketblocking.cc:126-391 does: voxels with E >= 15 T are
 * inaccessible, accessible voxels form 6-connected segments without periodic
 * images numbered in raster order of their first voxel, the faces x = N-1 ->
 * x = 0 (likewise y, z) link segments, and a segment whose part of that
 * graph holds a loop with non-zero net cell translation is a channel; every
 * other segment is a pocket.  With block_in_place, pocket voxels are set to
 * 1e13 (FindAndBlockPockets, src/pocketblocking.cc:373-382).
 * segments_out (nullable) receives the segment id of every voxel (-1 if
 * inaccessible), layout k + j*nz + i*ny*nz.
 * er_widom_grid scores n insertions of the Philox stream (seed, 1) by
 * trilinear interpolation of the grid (src/gcmc.cc:145-177 after the wrap
 * of src/gcmc.cc:74-77; lower index clamped to N-2). */
int er_grid_upload(er_context *ctx, const float *host, const int n[3], er_grid **out);
int er_pockets_label(er_context *ctx, er_grid *grid, double temperature, int block_in_place,
                     int *n_pockets, int32_t *segments_out);
/* Many grids labelled together (one launch per phase for all of them);
 * n_pockets[i] receives grid i's pocket count.  Same results as
 * er_pockets_label on each grid; no reference counterpart as a batch (the
 * reference labels one grid per FindAndBlockPockets call). */
int er_pockets_label_batch(er_context *ctx, er_grid *const *grids, int n_grids,
                           double temperature, int block_in_place, int *n_pockets);
int er_widom_grid(er_context *ctx, er_grid *grid, uint64_t n_insertions, double temperature,
                  uint64_t seed, er_widom_result *out, double *energies_out);
/* Guest-framework energies (K) at n_positions fractional positions
 * frac_xyz[3 n] of any range, by the wrap of src/gcmc.cc:74-77 and the
 * trilinear interpolation of GuestFrameworkEnergy (src/gcmc.cc:145-177) on a
 * resident FP32 or FP64 grid: the energy calls of GuestFrameworkTotalEnergy
 * (src/gcmc.cc:182-198) for many guests or chains in one call. */
int er_grid_energies(er_context *ctx, const er_grid *grid, const double *frac_xyz,
                     uint64_t n_positions, double *energies_out);

/* ---- serving ------------------------------------------------------------------
 * No reference counterpart (the reference runs one framework per process and
 * uploads it per run, src/writegrid.cu:186-189, src/henry.cu:133-135).  A
 * server owns a scheduler thread on one context: requests (a framework plus a
 * grid and/or a Widom run) are admitted in arrival order while the HBM budget
 * allows (the context's cap, or the device memory free at server creation
 * less a margin), and every launch of the tiled kernel takes pending tiles /
 * bins from all admitted requests, so work of requests that arrive while
 * others run overlaps on the device.  Frameworks stay resident in HBM: an
 * upload is shared by every request with the same atoms, cell, replication
 * and cutoff, and kept after its requests finish until admission needs the
 * memory (least recently used first) or the idle ones exceed cache_bytes.
 * Each launch computes FP64 grids or everything else, chosen by the oldest
 * ready request's pending work.  Results are bit-identical to the plain
 * path.  While a server runs, use its context only through the server. */
typedef struct er_server_config {
    int streams;               /* launches in flight (1..8) */
    int64_t items_per_launch;  /* tiles + bins per launch, 0 = all pending */
    int max_queue;             /* waiting requests beyond which submit rejects */
    int64_t cache_bytes;       /* idle frameworks kept resident: -1 up to the budget, 0 none */
    int fail_launch;           /* failure-path testing: the n-th launch attempt is refused
                                  on the host before anything is enqueued, and the requests
                                  holding its items fail with ER_E_HIP (0 = never) */
} er_server_config;

typedef struct er_request {
    const er_particle *atoms;
    int n_atoms;
    double t_matrix[9];
    int rep[3];
    double rc2;
    int grid_n[3];            /* all zero: no grid */
    int precision;
    uint64_t n_insertions;    /* 0: no Widom */
    double temperature;
    uint64_t seed;
} er_request;

typedef enum er_request_state {
    ER_RQ_QUEUED = 0, ER_RQ_RUNNING = 1, ER_RQ_DONE = 2, ER_RQ_FAILED = 3, ER_RQ_REJECTED = 4
} er_request_state;

typedef struct er_request_status {
    int state;                /* er_request_state */
    int status;               /* er_status of a failed or rejected request */
    double t_submit, t_admit, t_done;  /* seconds on the server's steady clock */
    er_widom_result widom;
} er_request_status;

int er_server_create(er_context *ctx, const er_server_config *config, er_server **out);
/* ER_E_CAPACITY (queue full) and ER_E_NOMEM (request alone exceeds the HBM
 * cap) are rejections: the ticket is still valid and reports REJECTED. */
int er_server_submit(er_server *server, const er_request *request, uint64_t *ticket);
/* wait_ms < 0 blocks until the request is done, failed or rejected. */
int er_server_poll(er_server *server, uint64_t ticket, double wait_ms, er_request_status *out);
int er_server_fetch_grid(er_server *server, uint64_t ticket, void *host_out);
int er_server_release(er_server *server, uint64_t ticket);
int er_server_stats(er_server *server, int64_t *submitted, int64_t *rejected, int64_t *completed,
                    int64_t *launches);
/* Framework residency: uploads avoided (hits), uploads made (misses),
 * frameworks evicted, and the frameworks / bytes resident now. */
int er_server_cache_stats(er_server *server, int64_t *hits, int64_t *misses, int64_t *evictions,
                          int64_t *resident_frameworks, int64_t *resident_bytes);
int er_server_destroy(er_server *server);

/* ---- host side of the drop-in binaries -------------------------------------
 * Everything src/writegrid.cu:64-127 / src/henry.cu:50-95 do before the
 * device work, from a directory laid out like the reference (simulation.input,
 * data/...).  On error the return code is set and er_last_error() holds the
 * message the reference prints. mode: 0 writegrid, 1 henry. */
typedef struct er_host_info {
    double a, b, c, alpha, beta, gamma; /* lengths (A), angles (radians) */
    int n_atoms;
    double density;                     /* kg/m^3 */
    double volume;                      /* A^3 */
    double t_matrix[9], inv_t_matrix[9];
    int rep[3];
    double r_cutoff_squared;
    double temperature;
    double epsilon_guest, sigma_guest, adsorbate_mw;
    int grid_n[3];                      /* writegrid */
    double grid_resolution;             /* writegrid */
    int grid_precision;                 /* writegrid: GridPrecision extension key */
    int insertions_per_a3;              /* henry */
    int numinsertions;                  /* henry: (int)(per_A3 * V) */
    int insertions_per_thread;          /* henry: numinsertions/16384 + 1 */
    int64_t total_insertions;           /* henry: 16384 * insertions_per_thread */
    uint64_t seed;                      /* henry: WidomSeed extension key, default 1234 */
} er_host_info;

int er_host_setup(const char *root, const char *structure, const char *adsorbate, int mode,
                  er_host_run **out);
int er_host_run_info(const er_host_run *run, er_host_info *out);
int er_host_run_particles(const er_host_run *run, er_particle *out, int capacity);
/* label of atom i (NUL-terminated, truncated to capacity) */
int er_host_run_label(const er_host_run *run, int i, char *out, int capacity);
int er_host_run_write_settings(const er_host_run *run, const char *path);
void er_host_run_release(er_host_run *run);

/* Grid file writer of src/writegrid.cu:144-171,244-269 ("txt" or "cube");
 * exactly one of e64 / e32 is non-null. */
int er_write_grid_file(const char *path, const char *format, const double *e64, const float *e32,
                       const int n[3], const double t_matrix[9], int threads);

/* Hard-sphere accessibility grid, ComputeGridSheetAccessibleOrNot
 * (src/computegridsheet.h:101-165): out[k + j*nz + i*ny*nz] = 1.0 where no
 * atom image with offsets in [-1, 1]^3 (whatever the replication factors)
 * lies strictly closer than sigma_ij (atoms[].sig2, mixed with the guest),
 * else 0.0; points f = i / (N - 1).  The reference sizes this grid as
 * N_i = ceil(1 / resolution) (src/writegrid.cu:104-108). */
int er_accessibility_grid(er_context *ctx, const er_particle *atoms, int n_atoms,
                          const double t_matrix[9], const int n[3], double *out);

/* ---- the reference's direct method on the GPU -------------------------------
 * ComputeGridSheet (src/computegridsheet.h:30-99, whole grid in one launch)
 * and doHenry (src/doHenry.h:29-135, one insertion per lane) with the
 * reference's arithmetic in FP64 over every (point, image, atom) triple in
 * its order, atoms in the caller's order: no pruning and no mixed precision.
 * A baseline for the tile kernel (the reference's method on the same GPU)
 * and a second device implementation of the reference arithmetic.  Grid:
 * FP64, layout and points as er_grid_compute, at most 2^31 points;
 * host_out and resident_out are each nullable.  Widom: the Philox points of
 * er_widom (same seed, same points), n_insertions in [1, 2^31], sums as
 * er_widom; energies_out (nullable) receives double[n_insertions]. */
int er_direct_grid(er_context *ctx, const er_particle *atoms, int n_atoms,
                   const double t_matrix[9], const int rep[3], double rc2, const int n[3],
                   double *host_out, er_grid **resident_out);
int er_direct_widom(er_context *ctx, const er_particle *atoms, int n_atoms,
                    const double t_matrix[9], const int rep[3], double rc2,
                    uint64_t n_insertions, double temperature, uint64_t seed,
                    er_widom_result *out, double *energies_out);
/* ---- grid files ---------------------------------------------------------
 * Text grids (src/writegrid.cu:144-147,260-269; the format er_write_grid_file
 * and the drop-in writegrid write) are read as src/gcmc.cc:286-346 imports
 * them: header "N_x N_y N_z ...", then N_x*N_y lines of N_z values (one
 * z-pencil per line), layout k + j*nz + i*ny*nz; a short or missing line is
 * an error.  No reference counterpart: a binary side-format (128-byte header
 * "ERGRID01", version, dtype, N_x, N_y, N_z, t_matrix; then raw
 * little-endian FP32 or FP64 values in the same layout) that the readers
 * recognise by its magic. */
int er_grid_file_info(const char *path, int n[3]);
int er_read_grid_file(const char *path, const int n_expected[3], double *out);
int er_write_grid_binary(const char *path, const double *e64, const float *e32, const int n[3],
                         const double t_matrix[9]);
/* A text or binary grid file as a resident FP32 grid (for pocket blocking and
 * grid-interpolated Widom). */
int er_grid_load(er_context *ctx, const char *path, er_grid **out);

/* Replication factors (image half-ranges) of julia_codes/framework.jl:162-211
 * for a row-major fractional->Cartesian matrix: per axis, 1 plus the number
 * of cell-vector doublings until the supercell centre is at least `cutoff`
 * (twice: 2 cutoff) from the face plane of the other two vectors. */
int er_replication_factors(const double t_matrix[9], double cutoff, int twice, int rep[3]);
/* Reads data/structures/<structure>.cssr under root and writes
 * data/uc_replications/<structure>_<once|twice>.uc ("%d %d %d", as
 * julia_codes/framework.jl:213-227); rep (nullable) receives the factors. */
int er_write_uc_file(const char *root, const char *structure, double cutoff, int twice,
                     int rep[3]);

#ifdef __cplusplus
}
#endif

#endif /* EMPTYROOM_H_ */
