// CPU-simulated launch of the particle-decomposition PIC program
// (progs/pic_pdecomp.hip):
//   run_pic_pdecomp PX PY NX NY NP CYCLES SEED RANKS THREADS
// Prints "energy C <FP64 bits>" per cycle, "particles N hash H" (FNV-1a of
// the final particles, float4 each) and "phases T0 .. T4".
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
    if (argc < 10) {
        std::fprintf(stderr, "usage: %s PX PY NX NY NP CYCLES SEED RANKS THREADS\n", argv[0]);
        return 2;
    }
    const int px = std::atoi(argv[1]), py = std::atoi(argv[2]), nx = std::atoi(argv[3]), ny = std::atoi(argv[4]);
    const long np = std::atol(argv[5]);
    const int cycles = std::atoi(argv[6]), P = std::atoi(argv[8]), threads = std::atoi(argv[9]);
    const uint64_t seed = std::strtoull(argv[7], nullptr, 10);
    const long n = (long)px * py * np;
    std::vector<double> e(cycles, 0.0), tp(5, -1.0);
    std::vector<float> g((size_t)n * 4, 0.0f);
    std::vector<uint64_t> words(64, 0);
    words[1] = px; words[2] = py; words[3] = nx; words[4] = ny; words[5] = np; words[6] = cycles; words[7] = seed;
    words[8] = fbits(0.02f); words[9] = fbits(0.01f); words[10] = fbits(0.001f);
    words[11] = (uint64_t)(uintptr_t)e.data();
    words[12] = (uint64_t)(uintptr_t)g.data();
    words[13] = (uint64_t)(uintptr_t)tp.data();
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    cfg.heap_bytes = (size_t)P * (1u << 20);
    gpumpi_result res;
    char name[] = "pic_pdecomp";
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup("pic_pdecomp"), 1, av, words.data(), &res);
    std::printf("pic_pdecomp P=%d rc=%d status=%d mpi_error=%d msg='%s'\n", P, rc, res.status, res.mpi_error,
                res.mpi_error_msg);
    for (int c = 0; c < cycles; c++) {
        uint64_t b;
        std::memcpy(&b, &e[c], 8);
        std::printf("energy %d %llu\n", c, (unsigned long long)b);
    }
    uint64_t h = 1469598103934665603ULL;
    const uint8_t* b = (const uint8_t*)g.data();
    for (size_t i = 0; i < g.size() * 4; i++) h = (h ^ b[i]) * 1099511628211ULL;
    std::printf("particles %ld hash %llu\n", n, (unsigned long long)h);
    std::printf("phases");
    for (double t : tp) std::printf(" %.9g", t);
    std::printf("\n");
    return rc;
}
