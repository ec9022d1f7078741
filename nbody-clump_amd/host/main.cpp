// nbody_clump: drop-in executable for the reference's main()
// (src/main.cpp:4-17). Runs the simulation from the current directory (or the
// directory given as the first argument) and exits with EXIT_FAILURE and the
// reference's message when the library reports a fatal condition. nbc.par's
// Driver key selects DEMmain's single-sphere loops instead of the clump loop.
// `--bin-to-text RUNDIR STEP OUTDIR` writes a run directory that resumes from
// the binary outputs of a step (nbc_resume_from_bin).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/nbc.h"

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "--bin-to-text") {
        // nbody_clump --bin-to-text RUNDIR STEP OUTDIR: a run directory that
        // resumes from the binary outputs of step STEP
        if (argc != 5) {
            std::fprintf(stderr, "usage: nbody_clump --bin-to-text RUNDIR STEP OUTDIR\n");
            return 2;
        }
        const int e = nbc_resume_from_bin(argv[2], std::atoi(argv[3]), argv[4]);
        if (e) std::fprintf(stderr, "%s\n", nbc_last_error(nullptr));
        return e ? 1 : 0;
    }
    const char* dir = argc > 1 ? argv[1] : ".";
    const int device = argc > 2 ? std::atoi(argv[2]) : 0;
    const int e = nbc_run_simulation(dir, device, 0);
    if (e) {
        std::fprintf(stderr, "%s\n", nbc_last_error(nullptr));
        return e == 7 ? 1 : EXIT_FAILURE;
    }
    std::cout << "Final!!" << std::endl;
    return 0;
}
