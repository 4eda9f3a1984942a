/*
 * er_oracle.c -- CPU FP64 restatement of EmptyRoom's hot-path arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (see er_oracle.h).  Compiled with
 * -ffp-contract=off so every expression rounds as written.
 */
#include "er_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */
/* ------------------------------------------------------------------ */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

static void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
    uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k[0] += PHILOX_W0;
            k[1] += PHILOX_W1;
        }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof(c));
}

void orc_widom_point(uint64_t seed, uint32_t stream, uint64_t i, double f[3]) {
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), stream, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    const double inv2_32 = 1.0 / 4294967296.0;
    for (int k = 0; k < 3; k++) f[k] = ((double)w[k] + 1.0) * inv2_32;
}

static void set_threads(int nthreads);

/* Points i0 .. i0 + n - 1 of one stream, f[3 n]. */
void orc_widom_points(uint64_t seed, uint32_t stream, uint64_t i0, int64_t n, double *f,
                      int nthreads) {
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) orc_widom_point(seed, stream, i0 + (uint64_t)p, f + 3 * p);
}

/* ------------------------------------------------------------------ */
/* Cell geometry and mixing                                            */
/* ------------------------------------------------------------------ */

void orc_cell_matrices(double a, double b, double c,
                       double alpha_deg, double beta_deg, double gamma_deg,
                       double t[9], double inv_t[9], double *volume) {
    /* src/Framework.cc:35-37 (degrees -> radians) */
    double al = alpha_deg * M_PI / 180.0;
    double be = beta_deg * M_PI / 180.0;
    double ga = gamma_deg * M_PI / 180.0;
    /* src/Framework.cc:113-114 */
    double v = sqrt(1.0 - pow(cos(al), 2.0) - pow(cos(be), 2) - pow(cos(ga), 2) +
                    2 * cos(al) * cos(be) * cos(ga));
    *volume = v * a * b * c;
    /* src/Framework.cc:124-132 */
    t[0] = a;
    t[1] = b * cos(ga);
    t[2] = c * cos(be);
    t[3] = 0.0;
    t[4] = b * sin(ga);
    t[5] = c * (cos(al) - cos(be) * cos(ga)) / sin(ga);
    t[6] = 0.0;
    t[7] = 0.0;
    t[8] = c * v / sin(ga);
    /* src/Framework.cc:134-142 */
    inv_t[0] = 1.0 / a;
    inv_t[1] = -cos(ga) / a / sin(ga);
    inv_t[2] = (cos(al) * cos(ga) - cos(be)) / (a * v * sin(ga));
    inv_t[3] = 0.0;
    inv_t[4] = 1.0 / b / sin(ga);
    inv_t[5] = (cos(be) * cos(ga) - cos(al)) / (b * v * sin(ga));
    inv_t[6] = 0.0;
    inv_t[7] = 0.0;
    inv_t[8] = sin(ga) / (c * v);
}

void orc_lb_mix(double eps_guest, double sig_guest, double eps_atom, double sig_atom,
                double *eps_ij, double *sig2_ij) {
    *sig2_ij = pow((sig_guest + sig_atom) / 2.0, 2);
    *eps_ij = sqrt(eps_guest * eps_atom);
}

/* ------------------------------------------------------------------ */
/* LJ energy, reference enumeration                                    */
/* ------------------------------------------------------------------ */

static inline void frac_to_cart(const double t[9], double xf, double yf, double zf,
                                double *x, double *y, double *z) {
    /* src/computegridsheet.h:21-28 */
    *x = t[0] * xf + t[1] * yf + t[2] * zf;
    *y = t[3] * xf + t[4] * yf + t[5] * zf;
    *z = t[6] * xf + t[7] * yf + t[8] * zf;
}

static int64_t n_images(const int rep[3]) {
    return (int64_t)(2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1);
}

void orc_image_atoms(const double t[9], const double *atoms, int n, const int rep[3],
                     double *out) {
    /* The image position T(f_a + n) depends only on (image, atom), so it is
     * evaluated once here with exactly the arithmetic of
     * src/computegridsheet.h:61-69 and reused for every point. */
    int64_t q = 0;
    for (int i = -rep[0]; i <= rep[0]; i++)
        for (int j = -rep[1]; j <= rep[1]; j++)
            for (int k = -rep[2]; k <= rep[2]; k++)
                for (int a = 0; a < n; a++) {
                    double xf = atoms[5 * a + 0] + 1.0 * i;
                    double yf = atoms[5 * a + 1] + 1.0 * j;
                    double zf = atoms[5 * a + 2] + 1.0 * k;
                    frac_to_cart(t, xf, yf, zf, &out[3 * q], &out[3 * q + 1], &out[3 * q + 2]);
                    q++;
                }
}

/* Sum over the precomputed image list; the loop order (images outer, atoms
 * inner) and every operation follow src/computegridsheet.h:71-82. */
static double energy_from_images(const double *img, const double *atoms, int n,
                                 int64_t nimg, double rc2, double min_r2,
                                 double x, double y, double z) {
    double energy = 0.0;
    int64_t q = 0;
    for (int64_t m = 0; m < nimg; m++) {
        for (int a = 0; a < n; a++, q++) {
            double dx = x - img[3 * q];
            double dy = y - img[3 * q + 1];
            double dz = z - img[3 * q + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            r2 = (r2 > min_r2) ? r2 : min_r2;
            if (r2 < rc2) {
                double s2 = atoms[5 * a + 4] / r2;
                double s6 = s2 * s2 * s2;
                energy += 4.0 * atoms[5 * a + 3] * s6 * (s6 - 1.0);
            }
        }
    }
    return energy;
}

static int64_t count_from_images(const double *img, int64_t nq, double rc2, double min_r2,
                                 double x, double y, double z) {
    int64_t c = 0;
    for (int64_t q = 0; q < nq; q++) {
        double dx = x - img[3 * q];
        double dy = y - img[3 * q + 1];
        double dz = z - img[3 * q + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        r2 = (r2 > min_r2) ? r2 : min_r2;
        c += (r2 < rc2);
    }
    return c;
}

static double *build_images(const double t[9], const double *atoms, int n, const int rep[3]) {
    int64_t nq = n_images(rep) * n;
    double *img = (double *)malloc((size_t)(nq > 0 ? nq : 1) * 3 * sizeof(double));
    if (img) orc_image_atoms(t, atoms, n, rep, img);
    return img;
}

static const double GRID_MIN_R = 1e-12;  /* src/computegridsheet.h:19 */
static const double HENRY_MIN_R = 1e-16; /* src/doHenry.h:15 */

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

double orc_lj_energy(const double t[9], const double *atoms, int n, const int rep[3],
                     double rc2, double min_r2, const double f[3]) {
    double *img = build_images(t, atoms, n, rep);
    if (!img) return NAN;
    double x, y, z;
    frac_to_cart(t, f[0], f[1], f[2], &x, &y, &z);
    double e = energy_from_images(img, atoms, n, n_images(rep), rc2, min_r2, x, y, z);
    free(img);
    return e;
}

void orc_lj_grid(const double t[9], const double *atoms, int n, const int rep[3],
                 double rc2, int nx, int ny, int nz,
                 int i_begin, int i_end, int i_step, double *out, int nthreads) {
    double *img = build_images(t, atoms, n, rep);
    if (!img) return;
    const double min_r2 = GRID_MIN_R * GRID_MIN_R;
    int64_t nimg = n_images(rep);
    int nsheets = 0;
    for (int i = i_begin; i < i_end; i += i_step) nsheets++;
    set_threads(nthreads);
    int64_t total = (int64_t)nsheets * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < total; p++) {
        int s = (int)(p / ((int64_t)ny * nz));
        int rem = (int)(p % ((int64_t)ny * nz));
        int j = rem / nz, k = rem % nz;
        int i = i_begin + s * i_step;
        /* src/writegrid.cu:122-127 */
        double xf = 1.0 * i / (nx - 1);
        double yf = 1.0 * j / (ny - 1);
        double zf = 1.0 * k / (nz - 1);
        double x, y, z;
        frac_to_cart(t, xf, yf, zf, &x, &y, &z);
        out[p] = energy_from_images(img, atoms, n, nimg, rc2, min_r2, x, y, z);
    }
    free(img);
}

void orc_lj_points(const double t[9], const double *atoms, int n, const int rep[3],
                   double rc2, double min_r2, const double *f, int64_t m,
                   double *out, int nthreads) {
    double *img = build_images(t, atoms, n, rep);
    if (!img) return;
    int64_t nimg = n_images(rep);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < m; p++) {
        double x, y, z;
        frac_to_cart(t, f[3 * p], f[3 * p + 1], f[3 * p + 2], &x, &y, &z);
        out[p] = energy_from_images(img, atoms, n, nimg, rc2, min_r2, x, y, z);
    }
    free(img);
}

/* src/computegridsheet.h:101-165: a point is accessible (1) unless some atom
   image with offsets in [-1, 1]^3 -- whatever the replication factors -- lies
   strictly closer than sigma_ij (r^2 < sig2, no clamp); grid points as the
   energy grid (f = i / (N - 1)). */
void orc_accessibility_grid(const double t[9], const double *atoms, int n, int nx, int ny,
                            int nz, double *out, int nthreads) {
    const int rep1[3] = {1, 1, 1};
    double *img = build_images(t, atoms, n, rep1);
    if (!img) return;
    const int64_t nimg = n_images(rep1);
    set_threads(nthreads);
    const int64_t total = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < total; p++) {
        const int i = (int)(p / ((int64_t)ny * nz));
        const int j = (int)((p / nz) % ny), k = (int)(p % nz);
        double x, y, z;
        frac_to_cart(t, 1.0 * i / (nx - 1), 1.0 * j / (ny - 1), 1.0 * k / (nz - 1), &x, &y, &z);
        int overlap = 0;
        int64_t q = 0;
        for (int64_t m = 0; m < nimg && !overlap; m++)
            for (int a = 0; a < n; a++, q++) {
                const double dx = x - img[3 * q], dy = y - img[3 * q + 1], dz = z - img[3 * q + 2];
                if (dx * dx + dy * dy + dz * dz < atoms[5 * a + 4]) overlap = 1;
            }
        out[p] = overlap ? 0.0 : 1.0;
    }
    free(img);
}

int64_t orc_count_incut_grid(const double t[9], const double *atoms, int n, const int rep[3],
                             double rc2, int nx, int ny, int nz,
                             int i_begin, int i_end, int i_step, int nthreads) {
    double *img = build_images(t, atoms, n, rep);
    if (!img) return -1;
    const double min_r2 = GRID_MIN_R * GRID_MIN_R;
    int64_t nq = n_images(rep) * n;
    int nsheets = 0;
    for (int i = i_begin; i < i_end; i += i_step) nsheets++;
    set_threads(nthreads);
    int64_t total = (int64_t)nsheets * ny * nz, count = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : count)
    for (int64_t p = 0; p < total; p++) {
        int s = (int)(p / ((int64_t)ny * nz));
        int rem = (int)(p % ((int64_t)ny * nz));
        int j = rem / nz, k = rem % nz;
        int i = i_begin + s * i_step;
        double x, y, z;
        frac_to_cart(t, 1.0 * i / (nx - 1), 1.0 * j / (ny - 1), 1.0 * k / (nz - 1), &x, &y, &z);
        count += count_from_images(img, nq, rc2, min_r2, x, y, z);
    }
    free(img);
    return count;
}

/* ------------------------------------------------------------------ */
/* Widom insertion                                                     */
/* ------------------------------------------------------------------ */

void orc_widom(const double t[9], const double *atoms, int n, const int rep[3],
               double rc2, double temperature, uint64_t seed,
               uint64_t i0, uint64_t i1, double *energies,
               double *sum_b, double *sum_eb, int nthreads) {
    int64_t m = (int64_t)(i1 - i0);
    double *e = energies ? energies : (double *)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double *img = build_images(t, atoms, n, rep);
    if (!img || !e) {
        *sum_b = NAN;
        *sum_eb = NAN;
        free(img);
        if (!energies) free(e);
        return;
    }
    const double min_r2 = HENRY_MIN_R * HENRY_MIN_R;
    int64_t nimg = n_images(rep);
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < m; p++) {
        double f[3], x, y, z;
        orc_widom_point(seed, 0u, i0 + (uint64_t)p, f);
        frac_to_cart(t, f[0], f[1], f[2], &x, &y, &z);
        e[p] = energy_from_images(img, atoms, n, nimg, rc2, min_r2, x, y, z);
    }
    /* src/doHenry.h:107 and :117-125: b = exp(-E/T); the two sums. */
    double sb = 0.0, seb = 0.0;
    for (int64_t p = 0; p < m; p++) {
        double b = exp(-e[p] / temperature);
        sb += b;
        seb += e[p] * b;
    }
    *sum_b = sb;
    *sum_eb = seb;
    free(img);
    if (!energies) free(e);
}

void orc_henry_counts(int insertions_per_a3, double volume, int *numinsertions,
                      int *per_thread, int64_t *total) {
    /* src/henry.cu:79 */
    int ni = (int)(insertions_per_a3 * volume);
    /* src/henry.cu:94-95 with NUM_BLOCKS*NUM_THREADS = 64*256 */
    int pt = ni / (64 * 256) + 1;
    *numinsertions = ni;
    *per_thread = pt;
    *total = (int64_t)pt * 64 * 256;
}

void orc_henry_result(double sum_b, double sum_eb, double n_total, double temperature,
                      double density, double *kh_mol_per_kg_pa, double *kh_mol_per_m3_pa,
                      double *e_avg_k, double *e_avg_kj_per_mol) {
    /* src/henry.cu:166-174 */
    double e_avg = sum_eb / sum_b;
    double kh = sum_b / n_total / temperature / 8.314;
    *e_avg_k = e_avg;
    *e_avg_kj_per_mol = e_avg * 8.314 / 1000.0;
    *kh_mol_per_m3_pa = kh;
    *kh_mol_per_kg_pa = kh / density;
}

/* ------------------------------------------------------------------ */
/* Pocket blocking                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
    int from, to, dir;
} conn_t;

static int64_t gidx(int i, int j, int k, int ny, int nz) {
    /* src/pocketblocking.cc:10-13 */
    return (int64_t)k + (int64_t)j * nz + (int64_t)i * ny * nz;
}

/* src/pocketblocking.cc:103-124: skip a connection already present with the
 * same (from, to) and, unless it is a self-connection, the same direction. */
static int conn_exists(const conn_t *c, int nc, int from, int to, int dir) {
    for (int q = 0; q < nc; q++)
        if (c[q].to == to && c[q].from == from && (to == from || dir == c[q].dir)) return 1;
    return 0;
}

int orc_find_and_block_pockets(double *grid, int nx, int ny, int nz, double temperature,
                               int32_t *segment_out) {
    int64_t nv = (int64_t)nx * ny * nz;
    int32_t *seg = (int32_t *)malloc((size_t)nv * sizeof(int32_t));
    int64_t *stack = (int64_t *)malloc((size_t)nv * sizeof(int64_t) + 8);
    if (!seg || !stack) {
        free(seg);
        free(stack);
        return -1;
    }
    /* src/pocketblocking.cc:130-145: E >= 15 T is inaccessible (-1). */
    double thr = 15.0 * temperature;
    const int32_t UNSEEN = -2;
    for (int64_t v = 0; v < nv; v++) seg[v] = (grid[v] >= thr) ? -1 : UNSEEN;

    /* src/pocketblocking.cc:152-251: 6-neighbour flood fill without periodic
     * images; segments numbered in raster order of their first voxel. */
    int nseg = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                int64_t g = gidx(i, j, k, ny, nz);
                if (seg[g] != UNSEEN) continue;
                int id = nseg++;
                int64_t sp = 0;
                seg[g] = id;
                stack[sp++] = g;
                while (sp > 0) {
                    int64_t v = stack[--sp];
                    int vi = (int)(v / ((int64_t)ny * nz));
                    int vj = (int)((v / nz) % ny);
                    int vk = (int)(v % nz);
                    int64_t nb[6];
                    int nn = 0;
                    if (vi != 0) nb[nn++] = gidx(vi - 1, vj, vk, ny, nz);
                    if (vi != nx - 1) nb[nn++] = gidx(vi + 1, vj, vk, ny, nz);
                    if (vj != 0) nb[nn++] = gidx(vi, vj - 1, vk, ny, nz);
                    if (vj != ny - 1) nb[nn++] = gidx(vi, vj + 1, vk, ny, nz);
                    if (vk != 0) nb[nn++] = gidx(vi, vj, vk - 1, ny, nz);
                    if (vk != nz - 1) nb[nn++] = gidx(vi, vj, vk + 1, ny, nz);
                    for (int q = 0; q < nn; q++)
                        if (seg[nb[q]] == UNSEEN) {
                            seg[nb[q]] = id;
                            stack[sp++] = nb[q];
                        }
                }
            }
    free(stack);

    /* src/pocketblocking.cc:273-321: connections across the three face pairs,
     * from the segment at index N-1 to the segment at index 0, direction +1. */
    int cap = 64, nc = 0;
    conn_t *conn = (conn_t *)malloc((size_t)cap * sizeof(conn_t));
    for (int face = 0; face < 3; face++) {
        int n1 = face == 0 ? ny : nx;
        int n2 = face == 2 ? ny : nz;
        for (int u = 0; u < n1; u++)
            for (int w = 0; w < n2; w++) {
                int64_t here, there;
                if (face == 0) {
                    here = gidx(0, u, w, ny, nz);
                    there = gidx(nx - 1, u, w, ny, nz);
                } else if (face == 1) {
                    here = gidx(u, 0, w, ny, nz);
                    there = gidx(u, ny - 1, w, ny, nz);
                } else {
                    here = gidx(u, w, 0, ny, nz);
                    there = gidx(u, w, nz - 1, ny, nz);
                }
                int s_here = seg[here], s_there = seg[there];
                if (s_here == -1 || s_there == -1) continue;
                int dir = face + 1;
                if (conn_exists(conn, nc, s_there, s_here, dir)) continue;
                if (nc == cap) {
                    cap *= 2;
                    conn = (conn_t *)realloc(conn, (size_t)cap * sizeof(conn_t));
                }
                conn[nc].from = s_there;
                conn[nc].to = s_here;
                conn[nc].dir = dir;
                nc++;
            }
    }

    /* src/pocketblocking.cc:39-100,345-365: walk the connection graph from
     * each segment carrying the unit-cell offset; reaching a segment again
     * with a different offset, or a self-connection, makes it a channel.
     * Whether such a loop exists is a property of the whole connected part
     * of the graph, so each part is walked once. */
    int *status = (int *)malloc((size_t)(nseg > 0 ? nseg : 1) * sizeof(int)); /* 0 unknown, 1 pocket, 2 channel */
    int *uc = (int *)malloc((size_t)(nseg > 0 ? nseg : 1) * 3 * sizeof(int));
    int *visited = (int *)calloc((size_t)(nseg > 0 ? nseg : 1), sizeof(int));
    char *edge_seen = (char *)calloc((size_t)(nc > 0 ? nc : 1), 1);
    int *deg = (int *)calloc((size_t)(nseg + 1), sizeof(int));
    int *adj = (int *)malloc((size_t)(2 * nc + 1) * sizeof(int));
    int *sstack = (int *)malloc((size_t)(nseg > 0 ? nseg : 1) * sizeof(int));
    int *members = (int *)malloc((size_t)(nseg > 0 ? nseg : 1) * sizeof(int));
    for (int q = 0; q < nc; q++) {
        deg[conn[q].from + 1]++;
        if (conn[q].to != conn[q].from) deg[conn[q].to + 1]++;
    }
    for (int s = 0; s < nseg; s++) deg[s + 1] += deg[s];
    {
        int *fill = (int *)calloc((size_t)(nseg > 0 ? nseg : 1), sizeof(int));
        for (int q = 0; q < nc; q++) {
            adj[deg[conn[q].from] + fill[conn[q].from]++] = q;
            if (conn[q].to != conn[q].from) adj[deg[conn[q].to] + fill[conn[q].to]++] = q;
        }
        free(fill);
    }
    for (int s = 0; s < nseg; s++) status[s] = 0;
    for (int s0 = 0; s0 < nseg; s0++) {
        if (status[s0] != 0) continue;
        int channel = 0, sp = 0, nm = 0;
        visited[s0] = 1;
        uc[3 * s0] = uc[3 * s0 + 1] = uc[3 * s0 + 2] = 0;
        sstack[sp++] = s0;
        while (sp > 0) {
            int u = sstack[--sp];
            members[nm++] = u;
            for (int e = deg[u]; e < deg[u + 1]; e++) {
                int q = adj[e];
                if (edge_seen[q]) continue;
                edge_seen[q] = 1;
                if (conn[q].from == conn[q].to) {
                    channel = 1;
                    continue;
                }
                int other = conn[q].to, dir = conn[q].dir;
                if (conn[q].to == u) {
                    other = conn[q].from;
                    dir = -dir;
                }
                int nu[3] = {uc[3 * u], uc[3 * u + 1], uc[3 * u + 2]};
                if (dir > 0) nu[dir - 1]++;
                else nu[-dir - 1]--;
                if (!visited[other]) {
                    visited[other] = 1;
                    uc[3 * other] = nu[0];
                    uc[3 * other + 1] = nu[1];
                    uc[3 * other + 2] = nu[2];
                    sstack[sp++] = other;
                } else if (nu[0] != uc[3 * other] || nu[1] != uc[3 * other + 1] ||
                           nu[2] != uc[3 * other + 2]) {
                    channel = 1;
                }
            }
        }
        for (int q = 0; q < nm; q++) status[members[q]] = channel ? 2 : 1;
    }
    int npockets = 0;
    for (int s = 0; s < nseg; s++) npockets += (status[s] == 1);
    /* src/pocketblocking.cc:373-382 */
    for (int64_t v = 0; v < nv; v++)
        if (seg[v] >= 0 && status[seg[v]] == 1) grid[v] = 1e13;
    if (segment_out) memcpy(segment_out, seg, (size_t)nv * sizeof(int32_t));
    free(seg);
    free(conn);
    free(status);
    free(uc);
    free(visited);
    free(edge_seen);
    free(deg);
    free(adj);
    free(sstack);
    free(members);
    return npockets;
}

/* ------------------------------------------------------------------ */
/* Trilinear interpolation                                             */
/* ------------------------------------------------------------------ */

double orc_wrap(double x, double z) { return x - z * floor(x / z); }

double orc_trilinear(const double *grid, int nx, int ny, int nz,
                     double xf, double yf, double zf) {
    double dx = 1.0 / (nx - 1), dy = 1.0 / (ny - 1), dz = 1.0 / (nz - 1);
    /* src/gcmc.cc:153-155, lower index clamped to N-2 */
    int ix = (int)floor(xf / dx);
    int iy = (int)floor(yf / dy);
    int iz = (int)floor(zf / dz);
    if (ix > nx - 2) ix = nx - 2;
    if (iy > ny - 2) iy = ny - 2;
    if (iz > nz - 2) iz = nz - 2;
    if (ix < 0) ix = 0;
    if (iy < 0) iy = 0;
    if (iz < 0) iz = 0;
    /* src/gcmc.cc:159-161 */
    double xd = (xf - ix * dx) / dx;
    double yd = (yf - iy * dy) / dy;
    double zd = (zf - iz * dz) / dz;
#define G(a, b, c) grid[gidx((a), (b), (c), ny, nz)]
    /* src/gcmc.cc:164-176 */
    double c00 = G(ix, iy, iz) * (1.0 - xd) + xd * G(ix + 1, iy, iz);
    double c10 = G(ix, iy + 1, iz) * (1.0 - xd) + xd * G(ix + 1, iy + 1, iz);
    double c01 = G(ix, iy, iz + 1) * (1.0 - xd) + xd * G(ix + 1, iy, iz + 1);
    double c11 = G(ix, iy + 1, iz + 1) * (1.0 - xd) + xd * G(ix + 1, iy + 1, iz + 1);
#undef G
    double c0 = c00 * (1.0 - yd) + c10 * yd;
    double c1 = c01 * (1.0 - yd) + c11 * yd;
    return c0 * (1 - zd) + c1 * zd;
}

void orc_widom_grid(const double *grid, int nx, int ny, int nz, double temperature,
                    uint64_t seed, uint64_t i0, uint64_t i1, double *energies,
                    double *sum_b, double *sum_eb, int nthreads) {
    int64_t m = (int64_t)(i1 - i0);
    double *e = energies ? energies : (double *)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    if (!e) {
        *sum_b = NAN;
        *sum_eb = NAN;
        return;
    }
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < m; p++) {
        double f[3];
        orc_widom_point(seed, 1u, i0 + (uint64_t)p, f);
        /* src/gcmc.cc:436-829 wrap every coordinate before the lookup. */
        e[p] = orc_trilinear(grid, nx, ny, nz, orc_wrap(f[0], 1.0), orc_wrap(f[1], 1.0),
                             orc_wrap(f[2], 1.0));
    }
    double sb = 0.0, seb = 0.0;
    for (int64_t p = 0; p < m; p++) {
        double b = exp(-e[p] / temperature);
        sb += b;
        seb += (b == 0.0) ? 0.0 : e[p] * b;
    }
    *sum_b = sb;
    *sum_eb = seb;
    if (!energies) free(e);
}
