// CPU-simulated launch of NPB IS (progs/npb_is.hip):
//   run_is CLASS P THREADS USER_LDS_BYTES
// Prints the verification count, the key total, full_verify errors and a
// checksum of the globally sorted keys.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s CLASS P THREADS USER_LDS\n", argv[0]);
        return 2;
    }
    const char cls = argv[1][0];
    const int P = std::atoi(argv[2]), threads = std::atoi(argv[3]), ulds = std::atoi(argv[4]);
    const int logk = cls == 'S' ? 16 : cls == 'W' ? 20 : cls == 'A' ? 23 : 25;
    std::vector<int64_t> out(8, 0);
    std::vector<double> tout(4, 0), ph(16, 0);
    std::vector<int32_t> keys((size_t)1 << logk, 0);
    std::vector<uint64_t> words(64, 0);
    words[1] = (uint64_t)(unsigned char)cls;
    words[2] = (uint64_t)(uintptr_t)out.data();
    words[3] = (uint64_t)(uintptr_t)tout.data();
    words[4] = (uint64_t)(uintptr_t)keys.data();
    words[5] = 10;
    words[6] = (uint64_t)(uintptr_t)ph.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.user_lds_bytes = ulds;
    cfg.max_requests = 64;
    cfg.heap_bytes = (size_t)P * (64u << 20) / (size_t)(P / 8 > 1 ? P / 8 : 1);
    cfg.timeout_s = 0;
    gpumpi_result res;
    char name[] = "npb_is";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("npb_is"), 1, av, words.data(), &res);
    uint64_t h = 1469598103934665603ull;
    for (int32_t k : keys) h = (h ^ (uint32_t)k) * 1099511628211ull;
    std::printf("npb_is class %c P=%d rc=%d status=%d mpi_error=%d passed=%lld total=%lld errors=%lld keyhash=%llu\n",
                cls, P, rc, res.status, res.mpi_error, (long long)out[0], (long long)out[1], (long long)out[2],
                (unsigned long long)h);
    return rc;
}
