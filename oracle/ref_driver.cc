// ref_driver.cc -- command-line driver around the EmptyRoom reference's own
// host C++ (compiled from the reference tree by oracle/Makefile into
// oracle/_ref/ref_host).  TEST INFRASTRUCTURE ONLY: it exists so the tests can
// compare the oracle and the product's host library against what the
// reference code itself computes.  The reference functions call exit() on
// bad input, so the tests run this as a subprocess.
//
// Modes (run from a directory laid out like the reference: simulation.input,
// data/structures, data/forcefields, data/masses.def, data/uc_replications):
//   ref_host grid  <structure> <adsorbate>   host half of src/writegrid.cu:52-137
//   ref_host henry <structure> <adsorbate>   host half of src/henry.cu:43-104
//   ref_host pockets <nx> <ny> <nz> <T> <in.f64> <out.f64>
//                                            src/pocketblocking.cc:126-391
// The first two print every loaded value with %.17g and write the reference's
// output_files/<fw>_<ads>_{grid,henry}.out header.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <vector>

#include "datatypes.h"
#include "Forcefield.h"
#include "Framework.h"
#include "readsettings.h"
#include "load_fast_particle_f_array.h"
#include "write_settings_to_outputfile.h"
#include "pocketblocking.h"

static void dump_framework(const Framework &fw) {
    printf("cell %.17g %.17g %.17g %.17g %.17g %.17g\n", fw.a, fw.b, fw.c, fw.alpha, fw.beta,
           fw.gamma);
    printf("noatoms %d\n", fw.noatoms);
    printf("density %.17g\n", fw.density);
    printf("volume %.17g\n", fw.volume_unitcell);
    printf("t_matrix");
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) printf(" %.17g", fw.t_matrix[i][j]);
    printf("\ninv_t_matrix");
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) printf(" %.17g", fw.inv_t_matrix[i][j]);
    printf("\n");
}

static void dump_particles(const Framework &fw, const FrameworkParticle *p) {
    for (int i = 0; i < fw.noatoms; i++)
        printf("atom %d %s %.17g %.17g %.17g %.17g %.17g %.17g\n", i, fw.identity[i].c_str(),
               p[i].x_f, p[i].y_f, p[i].z_f, fw.mass[i], p[i].eps_with_adsorbate,
               p[i].sig_with_adsorbate_squared);
}

static int run_grid(const char *structure, const char *adsorbate) {
    GridParameters parameters;
    parameters.frameworkname = structure;
    parameters.adsorbate = adsorbate;
    parameters.adsorbateMW = GetAdsorbateMW(parameters.adsorbate);
    ReadSimulationInputFile(parameters);
    TripleInt uc_reps = ReadUnitCellReplicationFile(parameters.frameworkname, "once");
    parameters.replication_factor_a = uc_reps.arg1;
    parameters.replication_factor_b = uc_reps.arg2;
    parameters.replication_factor_c = uc_reps.arg3;
    Forcefield forcefield(parameters.forcefieldname);
    Framework framework(parameters.frameworkname);
    parameters.N_framework_atoms = framework.noatoms;
    PairDouble eps_sig = GrabGuestForceFieldParams(forcefield, parameters.adsorbate);
    parameters.epsilon_guest = eps_sig.arg1;
    parameters.sigma_guest = eps_sig.arg2;
    std::vector<FrameworkParticle> atoms(framework.noatoms > 0 ? framework.noatoms : 1);
    LoadFastFrameworkParticleArray(atoms.data(), framework, forcefield, parameters.epsilon_guest,
                                   parameters.sigma_guest);
    parameters.N_x = static_cast<int>(ceil(framework.a / parameters.grid_resolution));
    parameters.N_y = static_cast<int>(ceil(framework.b / parameters.grid_resolution));
    parameters.N_z = static_cast<int>(ceil(framework.c / parameters.grid_resolution));
    printf("adsorbate_mw %.17g\n", parameters.adsorbateMW);
    printf("forcefield %s\n", parameters.forcefieldname.c_str());
    printf("grid_resolution %.17g\n", parameters.grid_resolution);
    printf("gridoutputformat %s\n", parameters.gridoutputformat.c_str());
    printf("r_cutoff_squared %.17g\n", parameters.r_cutoff_squared);
    printf("temperature %.17g\n", parameters.T);
    printf("energy_threshold %.17g\n", parameters.energy_threshold);
    printf("feynmanhibbs %d\n", (int)parameters.feynmanhibbs);
    printf("verbose %d\n", (int)parameters.verbose);
    printf("rep %d %d %d\n", parameters.replication_factor_a, parameters.replication_factor_b,
           parameters.replication_factor_c);
    printf("guest %.17g %.17g\n", parameters.epsilon_guest, parameters.sigma_guest);
    printf("grid_n %d %d %d\n", parameters.N_x, parameters.N_y, parameters.N_z);
    dump_framework(framework);
    dump_particles(framework, atoms.data());
    char outname[512];
    snprintf(outname, sizeof(outname), "output_files/%s_%s_grid.out",
             parameters.frameworkname.c_str(), parameters.adsorbate.c_str());
    FILE *out = fopen(outname, "w");
    if (!out) return 3;
    WriteSettingsToOutputfile(out, parameters, framework, forcefield, atoms.data());
    fclose(out);
    return 0;
}

static int run_henry(const char *structure, const char *adsorbate) {
    HenryParameters parameters;
    parameters.frameworkname = structure;
    parameters.adsorbate = adsorbate;
    parameters.adsorbateMW = GetAdsorbateMW(parameters.adsorbate);
    ReadSimulationInputFile(parameters);
    TripleInt uc_reps = ReadUnitCellReplicationFile(parameters.frameworkname, "once");
    parameters.replication_factor_a = uc_reps.arg1;
    parameters.replication_factor_b = uc_reps.arg2;
    parameters.replication_factor_c = uc_reps.arg3;
    Forcefield forcefield(parameters.forcefieldname);
    Framework framework(parameters.frameworkname);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) parameters.t_matrix[i][j] = framework.t_matrix[i][j];
    parameters.N_framework_atoms = framework.noatoms;
    parameters.numinsertions =
        static_cast<int>(parameters.numinsertionsperA3 * framework.volume_unitcell);
    PairDouble eps_sig = GrabGuestForceFieldParams(forcefield, parameters.adsorbate);
    parameters.epsilon_guest = eps_sig.arg1;
    parameters.sigma_guest = eps_sig.arg2;
    std::vector<FrameworkParticle> atoms(framework.noatoms > 0 ? framework.noatoms : 1);
    LoadFastFrameworkParticleArray(atoms.data(), framework, forcefield, parameters.epsilon_guest,
                                   parameters.sigma_guest);
    // launch geometry the reference fixes at 64 x 256 (src/doHenry.h:17-18)
    parameters.num_blocks = 64;
    parameters.num_threads = 256;
    parameters.numinsertionsperthread = parameters.numinsertions / (64 * 256) + 1;
    printf("adsorbate_mw %.17g\n", parameters.adsorbateMW);
    printf("forcefield %s\n", parameters.forcefieldname.c_str());
    printf("r_cutoff_squared %.17g\n", parameters.r_cutoff_squared);
    printf("temperature %.17g\n", parameters.T);
    printf("insertions_per_a3 %d\n", parameters.numinsertionsperA3);
    printf("numinsertions %d\n", parameters.numinsertions);
    printf("per_thread %d\n", parameters.numinsertionsperthread);
    printf("verbose %d\n", (int)parameters.verbose);
    printf("rep %d %d %d\n", parameters.replication_factor_a, parameters.replication_factor_b,
           parameters.replication_factor_c);
    printf("guest %.17g %.17g\n", parameters.epsilon_guest, parameters.sigma_guest);
    dump_framework(framework);
    dump_particles(framework, atoms.data());
    char outname[512];
    snprintf(outname, sizeof(outname), "output_files/%s_%s_henry.out",
             parameters.frameworkname.c_str(), parameters.adsorbate.c_str());
    FILE *out = fopen(outname, "w");
    if (!out) return 3;
    WriteSettingsToOutputfile(out, parameters, framework, forcefield, atoms.data());
    fclose(out);
    return 0;
}

static int run_pockets(int nx, int ny, int nz, double temperature, const char *in,
                       const char *outp) {
    size_t nv = (size_t)nx * ny * nz;
    std::vector<double> grid(nv);
    FILE *f = fopen(in, "rb");
    if (!f || fread(grid.data(), sizeof(double), nv, f) != nv) return 3;
    fclose(f);
    GridInfo info;
    info.N_x = nx;
    info.N_y = ny;
    info.N_z = nz;
    info.numtotalpts = (int)nv;
    info.dx_f = 1.0 / (nx - 1);
    info.dy_f = 1.0 / (ny - 1);
    info.dz_f = 1.0 / (nz - 1);
    GCMCParameters parameters;
    parameters.verbose = false;
    parameters.debugmode = false;
    int np = FindAndBlockPockets(grid.data(), info, temperature, parameters);
    f = fopen(outp, "wb");
    if (!f || fwrite(grid.data(), sizeof(double), nv, f) != nv) return 3;
    fclose(f);
    printf("npockets %d\n", np);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && std::string(argv[1]) == "grid") return run_grid(argv[2], argv[3]);
    if (argc == 4 && std::string(argv[1]) == "henry") return run_henry(argv[2], argv[3]);
    if (argc == 8 && std::string(argv[1]) == "pockets")
        return run_pockets(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atof(argv[5]), argv[6],
                           argv[7]);
    fprintf(stderr, "usage: ref_host grid|henry <structure> <adsorbate>\n"
                    "       ref_host pockets nx ny nz T in.f64 out.f64\n");
    return 2;
}
