// er_host.h -- host side of the drop-in writegrid / henry binaries: the input
// loaders, Lorentz-Berthelot mixing, grid sizing, Henry insertion counts and
// the text outputs, with the parsing rules, error messages and output formats
// of the EmptyRoom reference (citations per function).  All paths are
// relative to `root` (the reference uses the working directory).
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "er_runtime.h"

namespace er {
namespace host {

// src/Framework.h:13-34
struct Framework {
    std::string name;
    double a = 0, b = 0, c = 0;
    double alpha = 0, beta = 0, gamma = 0;  // radians
    int noatoms = 0;
    std::vector<double> x_f, y_f, z_f, mass;
    std::vector<std::string> identity;
    double density = 0, volume_unitcell = 0;
    double t_matrix[3][3] = {{0}}, inv_t_matrix[3][3] = {{0}};
};

// src/Forcefield.h:14-23
struct Forcefield {
    std::string name;
    int numinteractions = 0;
    std::vector<double> epsilon, sigma;
    std::vector<std::string> identity;
};

// Settings of a writegrid run (src/datatypes.h:78-110 minus the launch
// fields).  `precision` is an extension key (GridPrecision fp32|fp64) that the
// reference's scanner ignores.
struct GridParams {
    std::string forcefieldname = "None";
    bool verbose = false;
    double grid_resolution = -1.0;
    std::string gridoutputformat = "txt";
    double r_cutoff_squared = 12.5 * 12.5;
    bool feynmanhibbs = false;
    double energy_threshold = 0.0;
    double T = -1;
    std::string frameworkname, adsorbate;
    double epsilon_guest = 0, sigma_guest = 0, adsorbateMW = 0;
    int N_framework_atoms = 0;
    int N_x = 0, N_y = 0, N_z = 0;
    int rep[3] = {0, 0, 0};
    int precision = 1;  // 1 FP64 (reference arithmetic), 0 FP32 grid
    // Extension key AccessibilityGrid 0|1: the hard-sphere branch of
    // src/writegrid.cu:57-59,104-108,272-298, which the reference hardcodes off.
    bool accessibility = false;
};

// Settings of a henry run (src/datatypes.h:44-76).  `seed` is an extension
// key (WidomSeed) defaulting to the reference's MTGP32 seed 1234
// (src/henry.cu:126).
struct HenryParams {
    int numinsertions = 0;
    int numinsertionsperA3 = -1;
    int numinsertionsperthread = 0;
    std::string forcefieldname = "None";
    bool verbose = false;
    double r_cutoff_squared = 12.5 * 12.5;
    double T = -1;
    std::string frameworkname, adsorbate;
    double epsilon_guest = 0, sigma_guest = 0, adsorbateMW = 0;
    int N_framework_atoms = 0;
    int rep[3] = {0, 0, 0};
    int num_blocks = 64, num_threads = 256;  // src/doHenry.h:17-18
    unsigned long long seed = 1234;
};

// Each loader returns ST_OK or an error code with `msg` set to the exact text
// the reference prints before exit(EXIT_FAILURE).
int get_adsorbate_mw(const std::string &root, const std::string &adsorbate, double *mw,
                     std::string *msg);                                   // readsettings.h:19-49
int read_grid_settings(const std::string &root, GridParams *p, std::string *msg);   // :51-107
int read_henry_settings(const std::string &root, HenryParams *p, std::string *msg); // :109-156
int read_uc_file(const std::string &root, const std::string &framework,
                 const std::string &once_or_twice, int rep[3], std::string *msg);  // :243-274
int load_forcefield(const std::string &root, const std::string &name, Forcefield *ff,
                    std::string *msg);                                    // Forcefield.cc:15-53
int load_framework(const std::string &root, const std::string &name, Framework *fw,
                   std::string *msg);                                     // Framework.cc:13-178
int grab_guest_params(const Forcefield &ff, const std::string &adsorbate, double *eps,
                      double *sig, std::string *msg);                     // readsettings.h:276-294
int load_particles(const Framework &fw, const Forcefield &ff, double eps_guest, double sig_guest,
                   std::vector<Particle> *out, std::string *msg);  // load_fast_particle_f_array.h

// src/writegrid.cu:110-112
void grid_dims(const Framework &fw, double resolution, int n[3]);
// src/henry.cu:79,94-95
void henry_counts(const Framework &fw, HenryParams *p);
long long henry_total(const HenryParams &p);

// .out headers, src/write_settings_to_outputfile.h:11-101 and :103-189.
void write_grid_settings(FILE *f, const GridParams &p, const Framework &fw, const Forcefield &ff,
                         const Particle *atoms);
void write_henry_settings(FILE *f, const HenryParams &p, const Framework &fw,
                          const Forcefield &ff, const Particle *atoms);

// Grid files, src/writegrid.cu:144-171 (headers) and :244-269 (values).
// Values are formatted by `threads` workers into per-sheet buffers and written
// in order, byte-identical to the reference's fprintf loop.
void write_txt_header(FILE *f, const int n[3]);
void write_cube_header(FILE *f, const int n[3], const double t[3][3]);
int write_grid_values(FILE *f, const std::string &format, const double *e64, const float *e32,
                      const int n[3], int threads);

// Text grid (src/writegrid.cu:144-147,260-269) read as src/gcmc.cc:286-346
// does: a header line "N_x N_y N_z ...", then N_x*N_y lines, each the N_z
// values of one z-pencil (j fastest, then i); layout k + j*nz + i*ny*nz.
// Values beyond N_z on a line are ignored, as the reference's istream loop
// does; a short line, a missing line or a value that does not parse is an
// error (the reference would keep zeros or fail an assert).
int read_grid_txt(const std::string &path, int n[3], std::vector<double> *vals, int threads,
                  std::string *msg);
// Binary side-format: a 128-byte header {"ERGRID01", version 1, dtype
// (0 FP32, 1 FP64), N_x, N_y, N_z, 0, t_matrix[9] row-major, zero padding},
// then the values in the layout above, little-endian.
int write_grid_bin(const std::string &path, const int n[3], const double t[9], const double *e64,
                   const float *e32, std::string *msg);
int read_grid_bin(const std::string &path, int n[3], double t[9], int *dtype,
                  std::vector<char> *values, std::string *msg);
// Either format, by content (binary magic, else text), as FP64 values.
int read_grid_file(const std::string &path, int n[3], std::vector<double> *vals, int threads,
                   std::string *msg);

// Accessibility values of src/writegrid.cu:285-297: "%d " per point, a
// newline after every 6 values and after every z-pencil.
int write_accessibility_values(FILE *f, const double *acc, const int n[3]);

// Full host-side setup of a run, in the order of src/writegrid.cu:64-98 and
// src/henry.cu:50-92 (stops at the first error, like the reference).
struct RunSetup {
    Framework fw;
    Forcefield ff;
    std::vector<Particle> atoms;
};
int setup_grid_run(const std::string &root, const std::string &structure,
                   const std::string &adsorbate, GridParams *p, RunSetup *r, std::string *msg);
int setup_henry_run(const std::string &root, const std::string &structure,
                    const std::string &adsorbate, HenryParams *p, RunSetup *r, std::string *msg);

// Flattened fractional->Cartesian matrix.
void flat_t(const Framework &fw, double t[9]);

// Replication factors (julia_codes/framework.jl:162-211): starting from 1,
// an axis's factor grows while the distance from the centre of the current
// supercell to the face plane spanned by the other two cell vectors is below
// the cutoff ("once") or twice the cutoff ("twice"); each step doubles that
// cell vector (not (r+1)/r), so r = 1 + ceil(log2(2 r_c / w)) for a
// perpendicular width w < 2 r_c.  Returns ST_ARG for a degenerate cell.
int replication_factors(const double t[9], double cutoff, bool twice, int rep[3]);
// Cell of data/structures/<name>.cssr (first two lines; no masses file is
// needed, as in julia_codes/framework.jl:58-82), as a flattened matrix.
int read_cssr_cell(const std::string &root, const std::string &name, double t[9],
                   std::string *msg);
// data/uc_replications/<name>_<once|twice>.uc holding "%d %d %d"
// (julia_codes/framework.jl:213-227).
int write_uc_file(const std::string &root, const std::string &name, bool twice,
                  const int rep[3], std::string *msg);

}  // namespace host
}  // namespace er
