// writegrid -- drop-in replacement for the reference's writegrid binary
// (src/writegrid.cu:52-323): same command line, inputs, output files, .out
// header and grid formats; the grid is computed on the MI355X by
// libemptyroom (whole grid in one launch instead of one launch per x-sheet).
//
//   writegrid <structure> <adsorbate>
//
// simulation.input may carry the extra keys "GridPrecision fp32|fp64"
// (default fp64, the reference's arithmetic), "GridOutputFormat bin" (the
// binary side-format, data/grids/<fw>_<ads>_<ff>.grid) and "AccessibilityGrid 0|1"
// (the hard-sphere accessibility branch the reference hardcodes off,
// src/writegrid.cu:57); the reference's scanner ignores both.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "../../include/emptyroom.h"
#include "er_host.h"

using namespace er::host;

static double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static int gpu_fail(const char *what) {
    printf("Error in %s - %s \n", what, er_last_error());
    return EXIT_FAILURE;
}

int main(int argc, char *argv[]) {
    if (argc != 3) {
        printf("Run as:\n./writegrid structure_name AdsorbateID\n");
        exit(EXIT_FAILURE);
    }
    GridParams p;
    RunSetup r;
    std::string msg;
    if (setup_grid_run(".", argv[1], argv[2], &p, &r, &msg) != er::ST_OK) {
        printf("%s", msg.c_str());
        exit(EXIT_FAILURE);
    }
    if (p.accessibility) printf("\n\nPriting accessibility grid based on hard-sphere model...\n\n");
    if (p.verbose) printf("Read simulation.input, .uc file, forcefield and framework\n");

    char outputfilename[512];
    snprintf(outputfilename, sizeof(outputfilename), "output_files/%s_%s_grid.out",
             p.frameworkname.c_str(), p.adsorbate.c_str());
    FILE *outputfile = fopen(outputfilename, "w");
    if (!outputfile) {
        printf("Could not open %s\n", outputfilename);
        exit(EXIT_FAILURE);
    }
    write_grid_settings(outputfile, p, r.fw, r.ff, r.atoms.data());
    if (p.verbose) printf("Wrote info to outputfile\n");

    const int n[3] = {p.N_x, p.N_y, p.N_z};
    char gridfilename[512];
    FILE *gridfile = nullptr;
    if (p.accessibility && p.gridoutputformat != "cube") {
        // the reference opens no grid file for this combination
        // (src/writegrid.cu:144-156) and then writes through it
        printf("The accessibility grid is written in cube format only\n");
        exit(EXIT_FAILURE);
    }
    if (p.gridoutputformat == "txt") {
        snprintf(gridfilename, sizeof(gridfilename), "data/grids/%s_%s_%s.txt",
                 r.fw.name.c_str(), p.adsorbate.c_str(), r.ff.name.c_str());
        gridfile = fopen(gridfilename, "w");
        if (gridfile) write_txt_header(gridfile, n);
    } else if (p.gridoutputformat == "cube") {
        snprintf(gridfilename, sizeof(gridfilename),
                 p.accessibility ? "data/grids/%s_%s_%s_accessibility.cube"
                                 : "data/grids/%s_%s_%s.cube",
                 r.fw.name.c_str(), p.adsorbate.c_str(), r.ff.name.c_str());
        gridfile = fopen(gridfilename, "w");
        if (gridfile) write_cube_header(gridfile, n, r.fw.t_matrix);
    } else if (p.gridoutputformat == "bin") {  // extension: binary side-format
        snprintf(gridfilename, sizeof(gridfilename), "data/grids/%s_%s_%s.grid",
                 r.fw.name.c_str(), p.adsorbate.c_str(), r.ff.name.c_str());
        gridfile = fopen(gridfilename, "wb");
    } else {
        printf("Grid output format must be txt or cube\n");
        exit(EXIT_FAILURE);
    }
    if (!gridfile) {
        printf("Could not open %s\n", gridfilename);
        exit(EXIT_FAILURE);
    }
    // Informational lines of src/writegrid.cu:190,198,204, kept so the .out
    // reads the same; they describe the reference's FP64 buffers and launch.
    fprintf(outputfile, "    Size of framework atoms array = %f MB\n",
            r.fw.noatoms * 40 / (1024.0 * 1024.0));
    fprintf(outputfile, "    Size of grid sheet = %f MB\n",
            (double)p.N_z * p.N_y * 8 / (1024.0 * 1024.0));
    fprintf(outputfile, "    A block is %d by %d threads.\n", 16, 16);

    const double t0 = now_s();
    er_context *ctx = nullptr;
    er_framework *fw = nullptr;
    if (er_context_create(0, 0, &ctx) != ER_OK) return gpu_fail("er_context_create");
    double t[9];
    flat_t(r.fw, t);
    const size_t npts = (size_t)n[0] * n[1] * n[2];
    if (p.accessibility) {
        printf("Starting...\n");
        std::vector<double> acc(npts);
        if (er_accessibility_grid(ctx, reinterpret_cast<const er_particle *>(r.atoms.data()),
                                  r.fw.noatoms, t, n, acc.data()) != ER_OK)
            return gpu_fail("er_accessibility_grid");
        if (write_accessibility_values(gridfile, acc.data(), n) != er::ST_OK) {
            printf("Error writing %s\n", gridfilename);
            exit(EXIT_FAILURE);
        }
        fprintf(outputfile, "    Time to write grid: %f s\n", now_s() - t0);
        er_context_destroy(ctx);
        fclose(outputfile);
        fclose(gridfile);
        return 0;
    }
    if (er_framework_upload(ctx, reinterpret_cast<const er_particle *>(r.atoms.data()),
                            r.fw.noatoms, t, p.rep, p.r_cutoff_squared, &fw) != ER_OK)
        return gpu_fail("er_framework_upload");
    std::vector<double> e64;
    std::vector<float> e32;
    int st;
    if (p.precision == 1) {
        e64.resize(npts);
        st = er_grid_compute(ctx, fw, n, ER_FP64, e64.data(), nullptr);
    } else {
        e32.resize(npts);
        st = er_grid_compute(ctx, fw, n, ER_FP32, e32.data(), nullptr);
    }
    if (st != ER_OK) return gpu_fail("er_grid_compute");
    if (p.verbose) printf("Grid computed; writing %zu points\n", npts);
    const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (p.gridoutputformat == "bin") {
        fclose(gridfile);
        gridfile = nullptr;
        if (write_grid_bin(gridfilename, n, t, e64.empty() ? nullptr : e64.data(),
                           e32.empty() ? nullptr : e32.data(), &msg) != er::ST_OK) {
            printf("%s\n", msg.c_str());
            exit(EXIT_FAILURE);
        }
    } else if (write_grid_values(gridfile, p.gridoutputformat, e64.empty() ? nullptr : e64.data(),
                          e32.empty() ? nullptr : e32.data(), n, threads) != er::ST_OK) {
        printf("Error writing %s\n", gridfilename);
        exit(EXIT_FAILURE);
    }
    const double sim_time = now_s() - t0;
    fprintf(outputfile, "    Time to write grid: %f s\n", sim_time);
    if (p.verbose) printf("Completed grid writing!\n");
    er_framework_release(fw);
    er_context_destroy(ctx);
    fclose(outputfile);
    if (gridfile) fclose(gridfile);
    return 0;
}
