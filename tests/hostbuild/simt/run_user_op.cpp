// CPU-simulated launch of the user-defined operator program
// (progs/user_ops.hip):
//   run_user_op P COUNT ROOT IN_PLACE THREADS [HEAP_BYTES_PER_RANK]
// Prints "ok R V" per rank and "sum S" (the sum of the asymmetric result).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s P COUNT ROOT IN_PLACE THREADS\n", argv[0]);
        return 2;
    }
    const int P = std::atoi(argv[1]), n = std::atoi(argv[2]), root = std::atoi(argv[3]);
    const int in_place = std::atoi(argv[4]), threads = std::atoi(argv[5]);
    std::vector<int> ok(P, 0);
    std::vector<long long> out(n > 0 ? n : 1, 0);
    std::vector<uint64_t> words(64, 0);
    words[0] = (uint64_t)(uintptr_t)ok.data();
    words[1] = (uint64_t)n;
    words[2] = (uint64_t)(int64_t)root;
    words[3] = (uint64_t)in_place;
    words[4] = (uint64_t)(uintptr_t)out.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (argc > 6 ? std::strtoull(argv[6], nullptr, 10) : (1u << 20));
    gpumpi_result res;
    char name[] = "user_op_reduce";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("user_op_reduce"), 1, av, words.data(), &res);
    std::printf("user_op P=%d n=%d root=%d in_place=%d rc=%d status=%d mpi_error=%d msg='%s'\n", P, n, root, in_place,
                rc, res.status, res.mpi_error, res.mpi_error_msg);
    for (int r = 0; r < P; r++) std::printf("ok %d %d\n", r, ok[r]);
    long long s = 0;
    for (int i = 0; i < n; i++) s += out[i];
    std::printf("sum %lld\n", s);
    return rc;
}
