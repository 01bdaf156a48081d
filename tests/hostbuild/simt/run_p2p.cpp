// CPU-simulated launches of the p2p workloads (progs/p2p_bench.hip) under
// given launch knobs, with the received payloads checked on the host:
//   run_p2p ring|pingpong P BYTES THREADS EAGER_BYTES HELPER_MIN_BYTES CHAN_SLOTS
// Prints "mismatches N" (bytes that differ from the sender's payload).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

static uint8_t pattern(int rank, uint64_t i) { return (uint8_t)(rank * 37 + i * 11 + (i >> 8)); }

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s ring|pingpong P BYTES THREADS EAGER HELPER_MIN SLOTS\n", argv[0]);
        return 2;
    }
    const bool ring = std::strcmp(argv[1], "ring") == 0;
    const int P = std::atoi(argv[2]);
    const uint64_t bytes = std::strtoull(argv[3], nullptr, 10);
    const uint64_t stride = (bytes + 255) / 256 * 256 + 256;
    std::vector<uint8_t> sb(stride * P), rb(stride * P, 0);
    for (int r = 0; r < P; r++)
        for (uint64_t i = 0; i < bytes; i++) sb[r * stride + i] = pattern(r, i);
    uint64_t sizes[1] = {bytes};
    int64_t iters[1] = {3}, warm[1] = {1};
    std::vector<double> times(P, 0.0);
    std::vector<uint64_t> words(64, 0);
    words[1] = (uint64_t)(uintptr_t)sizes;
    words[2] = 1;
    words[3] = (uint64_t)(uintptr_t)iters;
    words[4] = (uint64_t)(uintptr_t)warm;
    words[5] = (uint64_t)(uintptr_t)sb.data();
    words[6] = (uint64_t)(uintptr_t)rb.data();
    words[7] = stride;
    words[8] = (uint64_t)(uintptr_t)times.data();
    words[9] = (uint64_t)(P / 2);
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = std::atoi(argv[4]);
    cfg.eager_bytes = std::atoi(argv[5]);
    cfg.helper_min_bytes = std::atoi(argv[6]);
    cfg.chan_slots = std::atoi(argv[7]);
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (1u << 20);
    gpumpi_result res;
    char name[] = "p2p";
    char* av[] = {name, nullptr};
    const char* prog = ring ? "p2p_ring_sweep" : "p2p_pingpong_sweep";
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup(prog), 1, av, words.data(), &res);
    std::printf("%s P=%d bytes=%llu rc=%d status=%d mpi_error=%d msg='%s'\n", prog, P, (unsigned long long)bytes, rc,
                res.status, res.mpi_error, res.mpi_error_msg);
    uint64_t bad = 0;
    for (int r = 0; r < P; r++) {
        const int src = ring ? (r + P - 1) % P : (r / 2 < P / 2 ? (r ^ 1) : -1);
        if (src < 0) continue;
        for (uint64_t i = 0; i < bytes; i++) bad += rb[r * stride + i] != pattern(src, i);
    }
    std::printf("mismatches %llu\n", (unsigned long long)bad);
    return rc;
}
