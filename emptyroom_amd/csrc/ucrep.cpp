// ucrep -- replication-factor tool: writes data/uc_replications/<s>_<once|twice>.uc
// for a structure, as julia_codes/framework.jl:213-227 (write_uc_replication_file)
// does, so the writegrid / henry drop-ins need no Julia.
//   ucrep <structure_name> <cutoff> <once|twice>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/emptyroom.h"

int main(int argc, char **argv) {
    if (argc != 4) {
        printf("Run as:\n./ucrep structure_name cutoff once|twice\n");
        return EXIT_FAILURE;
    }
    const std::string mode = argv[3];
    if (mode != "once" && mode != "twice") {
        printf("not valid for once_or_twice\n");  // julia_codes/framework.jl:219-221
        return EXIT_FAILURE;
    }
    char *end = nullptr;
    const double cutoff = strtod(argv[2], &end);
    if (!end || *end != '\0' || !(cutoff >= 0.0)) {
        printf("cutoff must be a non-negative number\n");
        return EXIT_FAILURE;
    }
    int rep[3];
    if (er_write_uc_file(".", argv[1], cutoff, mode == "twice", rep) != 0) {
        printf("%s\n", er_last_error());
        return EXIT_FAILURE;
    }
    printf("%s_%s.uc: %d %d %d\n", argv[1], mode.c_str(), rep[0], rep[1], rep[2]);
    return EXIT_SUCCESS;
}
