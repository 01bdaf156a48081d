// CPU-simulated launch of the particle-in-cell rank program (progs/pic.hip):
//   run_pic PX PY NX NY NP CYCLES SEED THREADS [PROGRAM]
// PROGRAM is `pic` (the default) or `pic_exact`.
// Prints "energy C <bits>" (FP64 bit patterns), "energy16 C <bits>",
// "phases T0 .. T6" (rank 0's phase seconds), "rank R count N phash H rhash H".
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

static uint64_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

int main(int argc, char** argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: %s PX PY NX NY NP CYCLES SEED THREADS\n", argv[0]);
        return 2;
    }
    const int px = std::atoi(argv[1]), py = std::atoi(argv[2]), nx = std::atoi(argv[3]), ny = std::atoi(argv[4]);
    const long np = std::atol(argv[5]);
    const int cycles = std::atoi(argv[6]), threads = std::atoi(argv[8]);
    const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
    const int P = px * py;
    const char* prog = argc > 9 ? argv[9] : "pic";
    std::vector<double> e(cycles, 0.0);
    std::vector<float> e16(cycles, 0.0f);
    std::vector<int64_t> counts(P, 0);
    std::vector<uint64_t> ph(P, 0), rh(P, 0);
    std::vector<double> tp(7, -1.0);
    std::vector<uint64_t> words(64, 0);
    words[1] = px; words[2] = py; words[3] = nx; words[4] = ny; words[5] = np; words[6] = cycles; words[7] = seed;
    words[8] = fbits(0.02f); words[9] = fbits(0.01f); words[10] = fbits(0.001f);
    words[11] = (uint64_t)(uintptr_t)e.data();
    words[12] = (uint64_t)(uintptr_t)e16.data();
    words[13] = (uint64_t)(uintptr_t)counts.data();
    words[14] = (uint64_t)(uintptr_t)ph.data();
    words[15] = (uint64_t)(uintptr_t)rh.data();
    words[16] = (uint64_t)(uintptr_t)tp.data();
    words[17] = (uint64_t)(np * 2 + 1024);
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (4u << 20);
    gpumpi_result res;
    char name[] = "pic";
    char* av[] = {name, nullptr};
    const void* kern = gpumpi_prog_lookup(prog);
    if (!kern) {
        std::fprintf(stderr, "no program %s\n", prog);
        return 2;
    }
    const int rc = gpumpi_run(&cfg, kern, 1, av, words.data(), &res);
    std::printf("pic P=%d rc=%d status=%d mpi_error=%d msg='%s'\n", P, rc, res.status, res.mpi_error, res.mpi_error_msg);
    for (int c = 0; c < cycles; c++) {
        uint64_t b;
        std::memcpy(&b, &e[c], 8);
        std::printf("energy %d %llu\nenergy16 %d %llu\n", c, (unsigned long long)b, c, (unsigned long long)fbits(e16[c]));
    }
    std::printf("phases");
    for (double t : tp) std::printf(" %.9g", t);
    std::printf("\n");
    for (int r = 0; r < P; r++)
        std::printf("rank %d count %lld phash %llu rhash %llu\n", r, (long long)counts[r], (unsigned long long)ph[r],
                    (unsigned long long)rh[r]);
    return rc;
}
