// plbfs <case.xml>: the reference's executable entry (src/main.cpp:18-28) on the
// MI355X solver.  Exit status 0 on success, 1 on any error.
#include <cstdio>

#include "plbfs_case.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <case.xml>\n", argv[0]);
        return 1;
    }
    plbfs_result r{};
    const int rc = plbfs_program_run(argv[1], &r);
    if (rc != PLBFS_OK) {
        std::fprintf(stderr, "plbfs: error %d: %s\n", rc, plbfs_last_error());
        return 1;
    }
    std::printf("steps %d converged %d residual rho %.6e u %.6e v %.6e w %.6e solve %.3f ms\n", r.total_steps,
                r.converged, r.residual[0], r.residual[1], r.residual[2], r.residual[3], r.solve_ms);
    return 0;
}
