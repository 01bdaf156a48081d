// CPU-simulated launch of the batched uneven Allreduce program
// (progs/batched.hip) on a layout file written by the test:
//   P C total, then P colours, C counts, P element offsets (text).
// usage: run_batched LAYOUT THREADS MAX_REQUESTS HEAP_PER_RANK ITERS WIRE...
// Prints one line per launch: wire, status, device error, per-colour
// relative errors against the program's FP64 reference.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);
extern "C" int gpumpi_fill_normal(void* p, uint64_t n, uint64_t seed, uint64_t stream);

int main(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s LAYOUT THREADS MAX_REQUESTS HEAP_PER_RANK ITERS WIRE...\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "r");
    if (!f) return 2;
    int P = 0, C = 0;
    unsigned long long total = 0;
    if (std::fscanf(f, "%d %d %llu", &P, &C, &total) != 3) return 2;
    std::vector<int> colors(P);
    std::vector<int64_t> counts(C);
    std::vector<uint64_t> offs(P);
    for (int& c : colors) if (std::fscanf(f, "%d", &c) != 1) return 2;
    for (int64_t& c : counts) if (std::fscanf(f, "%lld", (long long*)&c) != 1) return 2;
    for (uint64_t& o : offs) if (std::fscanf(f, "%llu", (unsigned long long*)&o) != 1) return 2;
    std::fclose(f);
    const int threads = std::atoi(argv[2]), R = std::atoi(argv[3]), iters = std::atoi(argv[5]);
    const size_t heap = (size_t)std::atoll(argv[4]);
    const void* kern = gpumpi_prog_lookup("batched_allreduce");
    if (!kern) return 3;

    // Payload buffers come from the (simulated) device allocator, as on the GPU.
    float *in = nullptr, *out = nullptr;
    double* ref = nullptr;
    if (gpumpi_device_alloc((void**)&in, 4 * total) || gpumpi_device_alloc((void**)&out, 4 * total) ||
        gpumpi_device_alloc((void**)&ref, 8 * total))
        return 4;
    for (int r = 0; r < P; r++)
        if (gpumpi_fill_normal(in + offs[r], (uint64_t)counts[colors[r]], 77, (uint64_t)r)) return 4;
    std::vector<double> times(P), errs(2 * (size_t)C);
    int bad = 0;
    for (int a = 6; a < argc; a++) {
        const int wire = std::atoi(argv[a]);
        std::fill(errs.begin(), errs.end(), 0.0);
        std::vector<uint64_t> words(64, 0);
        words[1] = (uint64_t)(uintptr_t)colors.data();
        words[2] = (uint64_t)(uintptr_t)counts.data();
        words[3] = (uint64_t)(uintptr_t)in;
        words[4] = (uint64_t)(uintptr_t)out;
        words[5] = (uint64_t)(uintptr_t)offs.data();
        words[6] = (uint64_t)wire;
        words[7] = (uint64_t)iters;
        words[8] = (uint64_t)(uintptr_t)times.data();
        words[9] = (uint64_t)(uintptr_t)errs.data();
        words[10] = (uint64_t)(uintptr_t)ref;
        gpumpi_config cfg;
        gpumpi_default_config(&cfg, P);
        cfg.threads_per_rank = threads;
        cfg.max_requests = R;
        cfg.heap_bytes = heap * (size_t)P;
        cfg.timeout_s = 0;
        gpumpi_result res;
        char name[] = "batched_allreduce";
        char* av[] = {name, nullptr};
        const int rc = gpumpi_run(&cfg, kern, 1, av, words.data(), &res);
        double worst = 0;
        for (int c = 0; c < C; c++) {
            const double e = std::sqrt(errs[2 * c] / errs[2 * c + 1]);
            if (std::getenv("SIMT_VERBOSE")) std::printf("  colour %d: err2 %.6e ref2 %.6e\n", c, errs[2 * c], errs[2 * c + 1]);
            if (!(e <= worst)) worst = e;  // NaN (no report) propagates
        }
        std::printf("wire=%d rc=%d status=%d mpi_error=%d rank=%d line=%d msg='%s' max_rel_err=%.3e lds=%d\n", wire, rc,
                    res.status, res.mpi_error, res.mpi_error_rank, res.mpi_error_line, res.mpi_error_msg, worst,
                    res.lds_bytes);
        std::fflush(stdout);
        if (rc) {
            std::printf("error: %s\n", gpumpi_last_error());
            bad = 1;
        }
        // Every member's output equals the rank-order FP32 fold of the
        // members' inputs for the exact path.
        if (wire == 0) {
            long long mism = 0;
            for (int c = 0; c < C; c++) {
                std::vector<int> mem;
                for (int r = 0; r < P; r++)
                    if (colors[r] == c) mem.push_back(r);
                const int64_t n = counts[c];
                for (int64_t i = 0; i < n; i++) {
                    float acc = in[offs[mem[0]] + i];
                    for (size_t k = 1; k < mem.size(); k++) acc = in[offs[mem[k]] + i] + acc;
                    for (int r : mem) mism += std::memcmp(&out[offs[r] + i], &acc, 4) != 0;
                }
            }
            std::printf("wire=0 output mismatches vs rank-order fold: %lld\n", mism);
            if (mism) bad = 1;
        }
        if (!(worst <= (wire == 0 ? 1e-6 : wire == 1 ? 1e-2 : 6e-2))) bad = 1;
    }
    return bad;
}
