/*
 * er_oracle.h -- CPU FP64 restatement of EmptyRoom's grid / Widom / pocket
 * arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in emptyroom_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only to check or to time
 * the CPU baseline; the product never links, loads or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the EmptyRoom source tree).  How each one is pinned is listed in
 * tests/test_oracle_*.py; a function without an independent pin says
 * "parity unpinned" here.
 */
#ifndef ER_ORACLE_H_
#define ER_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11).
 * Pinned by the Random123 known-answer vectors (tests/test_oracle_philox.py). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Insertion point `i` of Widom stream (`seed`, `stream`): three fractional
 * coordinates in (0,1], f_k = (w_k + 1) * 2^-32 from one Philox block with
 * counter (lo(i), hi(i), stream, 0) and key (lo(seed), hi(seed)).
 * Replaces curand_uniform_double on MTGP32 (src/doHenry.h:53-55), whose
 * stream is closed and is not reproduced. */
void orc_widom_point(uint64_t seed, uint32_t stream, uint64_t i, double f[3]);
void orc_widom_points(uint64_t seed, uint32_t stream, uint64_t i0, int64_t n, double *f,
                      int nthreads);

/* Fractional->Cartesian matrix, its inverse and the cell volume, as
 * src/Framework.cc:113-142 (angles in degrees, converted as :35-37). */
void orc_cell_matrices(double a, double b, double c,
                       double alpha_deg, double beta_deg, double gamma_deg,
                       double t[9], double inv_t[9], double *volume);

/* Lorentz-Berthelot mixing, src/load_fast_particle_f_array.h:35-36. */
void orc_lb_mix(double eps_guest, double sig_guest, double eps_atom, double sig_atom,
                double *eps_ij, double *sig2_ij);

/* Cartesian positions of every (image, atom) pair in the reference's
 * enumeration order: images i=-ra..ra (outer), j, k (inner), then atoms
 * (src/computegridsheet.h:54-69).  out: (2ra+1)(2rb+1)(2rc+1)*n*3 doubles. */
void orc_image_atoms(const double t[9], const double *atoms, int n, const int rep[3],
                     double *out);

/* Guest-framework LJ energy at fractional point f (src/computegridsheet.h:50-92,
 * src/doHenry.h:58-101).  atoms: n x 5 doubles (x_f, y_f, z_f, eps_ij, sig2_ij)
 * as FrameworkParticle (src/datatypes.h:20-28).  min_r2 is the r^2 clamp:
 * 1e-12^2 on the grid path (computegridsheet.h:19), 1e-16^2 on the Henry
 * path (doHenry.h:15). */
double orc_lj_energy(const double t[9], const double *atoms, int n, const int rep[3],
                     double rc2, double min_r2, const double f[3]);

/* Energy grid sheets i = i_begin, i_begin+i_step, ... < i_end with grid
 * points f = (i/(nx-1), j/(ny-1), k/(nz-1)) (src/writegrid.cu:103-127).
 * out holds one ny*nz sheet per computed i, index k + j*nz
 * (src/computegridsheet.h:94-97).  Uses the grid clamp 1e-24. */
void orc_lj_grid(const double t[9], const double *atoms, int n, const int rep[3],
                 double rc2, int nx, int ny, int nz,
                 int i_begin, int i_end, int i_step, double *out, int nthreads);

/* Energies at m arbitrary fractional points (m x 3). */
/* Hard-sphere accessibility grid, src/computegridsheet.h:101-165 (1 / 0). */
void orc_accessibility_grid(const double t[9], const double *atoms, int n, int nx, int ny,
                            int nz, double *out, int nthreads);
void orc_lj_points(const double t[9], const double *atoms, int n, const int rep[3],
                   double rc2, double min_r2, const double *f, int64_t m,
                   double *out, int nthreads);

/* Number of (point, atom, image) triples with r^2 < rc^2 over the same sheets
 * as orc_lj_grid: the unit of the FP32 roofline figure. */
int64_t orc_count_incut_grid(const double t[9], const double *atoms, int n, const int rep[3],
                             double rc2, int nx, int ny, int nz,
                             int i_begin, int i_end, int i_step, int nthreads);

/* Widom insertions i in [i0, i1) of stream (seed, 0): per-insertion energy
 * with the Henry clamp 1e-32, Boltzmann factor exp(-E/T), and the two FP64
 * sums in index order (src/doHenry.h:49-127).  energies may be NULL. */
void orc_widom(const double t[9], const double *atoms, int n, const int rep[3],
               double rc2, double temperature, uint64_t seed,
               uint64_t i0, uint64_t i1, double *energies,
               double *sum_b, double *sum_eb, int nthreads);

/* Insertion count arithmetic of src/henry.cu:79,94-95 (16384 = 64 blocks x
 * 256 threads, src/doHenry.h:17-18). */
void orc_henry_counts(int insertions_per_a3, double volume, int *numinsertions,
                      int *per_thread, int64_t *total);

/* K_H and <E> from the Boltzmann sums, src/henry.cu:157-174. */
void orc_henry_result(double sum_b, double sum_eb, double n_total, double temperature,
                      double density, double *kh_mol_per_kg_pa, double *kh_mol_per_m3_pa,
                      double *e_avg_k, double *e_avg_kj_per_mol);

/* Pocket blocking, src/pocketblocking.cc:126-391.  Voxels with E >= 15 T are
 * inaccessible (segment -1); accessible voxels form 6-connected segments
 * without periodicity, numbered by first raster visit; faces x=0/x=N-1 etc.
 * give directed connections; a segment whose connection graph holds a loop
 * with non-zero net cell translation is a channel, every other segment is a
 * pocket and its voxels are overwritten with 1e13.  segment_out (nullable)
 * receives the segment id of every voxel.  Returns the number of pockets. */
int orc_find_and_block_pockets(double *grid, int nx, int ny, int nz, double temperature,
                               int32_t *segment_out);

/* x - z*floor(x/z), src/gcmc.cc:74-77. */
double orc_wrap(double x, double z);

/* Trilinear interpolation of an endpoint-inclusive grid at fractional point
 * (xf, yf, zf) already in [0,1], src/gcmc.cc:145-177 with dx_f = 1/(N-1)
 * (src/gcmc.cc:343-345).  The lower index is clamped to N-2 so that f = 1
 * reads node N-1 instead of one past the grid (the reference would read out
 * of bounds there). */
double orc_trilinear(const double *grid, int nx, int ny, int nz,
                     double xf, double yf, double zf);

/* Widom insertions i in [i0, i1) of stream (seed, 1) scored by trilinear
 * interpolation of `grid` (after pocket blocking, if any). */
void orc_widom_grid(const double *grid, int nx, int ny, int nz, double temperature,
                    uint64_t seed, uint64_t i0, uint64_t i1, double *energies,
                    double *sum_b, double *sum_eb, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* ER_ORACLE_H_ */
