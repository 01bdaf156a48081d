// CPU-simulated launch of the MPI-IO and printf scenarios (progs/io.hip):
//   run_io P DIR N
// Prints "rank R: v0 v1 ..." (the rank's records) and "text R: <sprintf>".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s P DIR N\n", argv[0]);
        return 2;
    }
    const int P = std::atoi(argv[1]), N = std::atoi(argv[3]);
    const int kRec = 64;
    std::vector<int> rec((size_t)P * kRec, 0);
    std::vector<char> txt((size_t)P * 256, 0);
    std::vector<char> dir(argv[2], argv[2] + std::strlen(argv[2]) + 1);
    std::vector<uint64_t> words(64, 0);
    words[0] = (uint64_t)(uintptr_t)rec.data();
    words[1] = (uint64_t)(uintptr_t)dir.data();
    words[2] = (uint64_t)(uintptr_t)txt.data();
    words[3] = (uint64_t)N;
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = 128;
    cfg.timeout_s = 0;
    gpumpi_result res;
    char name[] = "io_scenarios";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("io_scenarios"), 1, av, words.data(), &res);
    std::printf("io_scenarios P=%d rc=%d status=%d mpi_error=%d msg='%s'\n", P, rc, res.status, res.mpi_error,
                res.mpi_error_msg);
    for (int r = 0; r < P; r++) {
        std::printf("rank %d:", r);
        for (int k = 0; k < kRec; k++) std::printf(" %d", rec[(size_t)r * kRec + k]);
        std::printf("\ntext %d: %s\n", r, &txt[(size_t)r * 256]);
    }
    return rc;
}
