// CPU-simulated launch of the resource-lifetime program (progs/lifetime.hip):
//   run_lifetime MODE P ITERATIONS THREADS [COMM_POOL_BYTES]
// Prints "ok R V" per rank and "out A B C D".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s MODE P ITERATIONS THREADS [COMM_POOL_BYTES]\n", argv[0]);
        return 2;
    }
    const int mode = std::atoi(argv[1]), P = std::atoi(argv[2]), iters = std::atoi(argv[3]);
    const int threads = std::atoi(argv[4]);
    std::vector<int> ok(P, 0);
    std::vector<int64_t> out(4, 0);
    std::vector<uint64_t> words(64, 0);
    words[0] = (uint64_t)(uintptr_t)ok.data();
    words[1] = (uint64_t)mode;
    words[2] = (uint64_t)iters;
    words[3] = (uint64_t)(uintptr_t)out.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (2u << 20);  // the arena stays below the size the simulation treats as bulk data
    if (argc > 5) cfg.comm_pool_bytes = std::strtoull(argv[5], nullptr, 10);
    gpumpi_result res;
    char name[] = "lifetime";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("lifetime"), 1, av, words.data(), &res);
    std::printf("lifetime mode=%d P=%d rc=%d status=%d mpi_error=%d msg='%s'\n", mode, P, rc, res.status,
                res.mpi_error, res.mpi_error_msg);
    for (int r = 0; r < P; r++) std::printf("ok %d %d\n", r, ok[r]);
    std::printf("out %lld %lld %lld %lld\n", (long long)out[0], (long long)out[1], (long long)out[2],
                (long long)out[3]);
    return rc;
}
