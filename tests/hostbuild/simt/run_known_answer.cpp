// CPU-simulated launch of one known-answer rank program
// (progs/known_answer.hip, args[0] = int ok[P]):
//   run_known_answer NAME P THREADS [MAX_REQUESTS]
// Prints "ok R V" per rank.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s NAME P THREADS [MAX_REQUESTS]\n", argv[0]);
        return 2;
    }
    const char* name = argv[1];
    const int P = std::atoi(argv[2]), threads = std::atoi(argv[3]);
    const void* prog = gpumpi_prog_lookup(name);
    if (!prog) {
        std::fprintf(stderr, "no program named %s\n", name);
        return 2;
    }
    std::vector<int> ok(P, 0);
    std::vector<uint64_t> words(64, 0);
    words[0] = (uint64_t)(uintptr_t)ok.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    if (argc > 4) cfg.max_requests = std::atoi(argv[4]);
    gpumpi_result res;
    char* av[] = {argv[1], nullptr};
    const int rc = gpumpi_run(&cfg, prog, 1, av, words.data(), &res);
    std::printf("%s P=%d threads=%d rc=%d status=%d mpi_error=%d msg='%s'\n", name, P, threads, rc, res.status,
                res.mpi_error, res.mpi_error_msg);
    for (int r = 0; r < P; r++) std::printf("ok %d %d\n", r, ok[r]);
    return rc;
}
