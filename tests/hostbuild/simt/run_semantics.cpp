// CPU-simulated launch of the semantics scenarios (progs/semantics.hip):
//   run_semantics P THREADS MASK [HEAP_BYTES_PER_RANK]
// The rank heaps together stay under the simulation's bulk-allocation size
// (8 MiB, simt.cpp), so that the lanes' accesses to heap buffers are in
// lockstep as on the GPU: the scenarios write a receive buffer from every
// lane just before the receive fills it.
// Prints one line per rank: "rank R: v0 v1 ..." (the rank's record).
//   run_semantics truncate RENDEZVOUS
// runs sem_truncate on 2 ranks and prints the launch's error record.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

static int run_truncate(int rendezvous) {
    std::vector<uint64_t> words(64, 0);
    words[1] = (uint64_t)rendezvous;
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, 2);
    cfg.threads_per_rank = 64;
    cfg.timeout_s = 0;
    cfg.heap_bytes = 2u << 20;
    gpumpi_result res;
    char name[] = "sem_truncate";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("sem_truncate"), 1, av, words.data(), &res);
    std::printf("sem_truncate rendezvous=%d rc=%d status=%d mpi_error=%d mpi_error_rank=%d msg='%s'\n", rendezvous,
                rc, res.status, res.mpi_error, res.mpi_error_rank, res.mpi_error_msg);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 3 && std::string(argv[1]) == "truncate") return run_truncate(std::atoi(argv[2]));
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s P THREADS MASK\n", argv[0]);
        return 2;
    }
    const int P = std::atoi(argv[1]), threads = std::atoi(argv[2]), mask = std::atoi(argv[3]);
    const int cap = 512;
    std::vector<int> rec((size_t)P * cap, 0), lens(P, 0);
    std::vector<uint64_t> words(64, 0);
    words[0] = (uint64_t)(uintptr_t)rec.data();
    words[1] = (uint64_t)cap;
    words[2] = (uint64_t)mask;
    words[3] = (uint64_t)(uintptr_t)lens.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (argc > 4 ? std::strtoull(argv[4], nullptr, 10) : (512u << 10));
    gpumpi_result res;
    char name[] = "semantics";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("semantics"), 1, av, words.data(), &res);
    std::printf("semantics P=%d rc=%d status=%d mpi_error=%d msg='%s'\n", P, rc, res.status, res.mpi_error,
                res.mpi_error_msg);
    for (int r = 0; r < P; r++) {
        std::printf("rank %d:", r);
        for (int i = 0; i < lens[r] && i < cap; i++) std::printf(" %d", rec[(size_t)r * cap + i]);
        std::printf("\n");
    }
    return rc;
}
