// henry -- drop-in replacement for the reference's henry binary
// (src/henry.cu:43-175): same command line, inputs, insertion-count
// arithmetic (16384 * (n/16384 + 1) insertions) and .out report; the Widom
// insertions run on the MI355X by libemptyroom.
//
//   henry <structure> <adsorbate>
//
// Insertion points come from a Philox4x32-10 stream keyed by the seed
// (extension key "WidomSeed", default 1234, the reference's MTGP32 seed); the
// reference's cuRAND MTGP32 stream is not reproduced, so K_H agrees with the
// reference statistically, not digit for digit.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

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
        printf("Run as:\n./henry structure_name AdsorbateID\n");
        exit(EXIT_FAILURE);
    }
    HenryParams p;
    RunSetup r;
    std::string msg;
    if (setup_henry_run(".", argv[1], argv[2], &p, &r, &msg) != er::ST_OK) {
        printf("%s", msg.c_str());
        exit(EXIT_FAILURE);
    }
    if (p.verbose) printf("# blocks = %d, # threads = %d\n", p.num_blocks, p.num_threads);

    char outputfilename[512];
    snprintf(outputfilename, sizeof(outputfilename), "output_files/%s_%s_henry.out",
             p.frameworkname.c_str(), p.adsorbate.c_str());
    FILE *outputfile = fopen(outputfilename, "w");
    if (!outputfile) {
        printf("Could not open %s\n", outputfilename);
        exit(EXIT_FAILURE);
    }
    write_henry_settings(outputfile, p, r.fw, r.ff, r.atoms.data());
    if (p.verbose) printf("Wrote info to outputfile\n");

    er_context *ctx = nullptr;
    er_framework *fw = nullptr;
    if (er_context_create(0, 0, &ctx) != ER_OK) return gpu_fail("er_context_create");
    double t[9];
    flat_t(r.fw, t);
    if (er_framework_upload(ctx, reinterpret_cast<const er_particle *>(r.atoms.data()),
                            r.fw.noatoms, t, p.rep, p.r_cutoff_squared, &fw) != ER_OK)
        return gpu_fail("er_framework_upload");
    const long long total = henry_total(p);
    const double start = now_s();
    er_widom_result res;
    if (er_widom(ctx, fw, (uint64_t)total, p.T, p.seed, &res, nullptr) != ER_OK)
        return gpu_fail("er_widom");
    const double sim_time = now_s() - start;
    fprintf(outputfile, "\nHenry simulation results\n");
    fprintf(outputfile, "    Simulation time: %f s\n", sim_time);
    if (p.verbose) printf("Copied data back from device to host\n");
    // src/henry.cu:166-174
    const double ensemble_average_energy = res.sum_eb / res.sum_b;
    const double henry_coeff = res.sum_b / (double)total / p.T / 8.314;
    fprintf(outputfile, "    Temperature (K): %f\n", p.T);
    fprintf(outputfile, "    <energy> (kJ/mol): %f\n", ensemble_average_energy * 8.314 / 1000.0);
    fprintf(outputfile, "    <energy> (K): %f\n", ensemble_average_energy);
    fprintf(outputfile, "    Henry coefficient (mol/(kg-Pa)): %e\n", henry_coeff / r.fw.density);
    fprintf(outputfile, "    Henry coefficient (mol/(m3-Pa)): %e\n", henry_coeff);
    fclose(outputfile);
    er_framework_release(fw);
    er_context_destroy(ctx);
    return 0;
}
