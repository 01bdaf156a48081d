// CPU-simulated launches of the collective test programs (progs/coll_bench.hip)
// with the shapes of the GPU tests, checked against the expected layouts.
//   run_coll alltoallv COUNTS_FILE THREADS     (COUNTS_FILE: P, then P*P counts)
//   run_coll blocks KIND THREADS               (P = 6, block 5, root 2)
//   run_coll abort P THREADS                   (rank 1 aborts; the others keep
//                                               calling collectives)
//   run_coll wire FMT P N THREADS              (compressed-wire Allreduce)
//   run_coll truncate THREADS                  (Alltoall send block larger
//                                               than the receive block)
//   run_coll race_selftest ORDERED             (the happens-before checker
//                                               on a planted race, SIMT_RACE=1)
#include "progs/registry.hpp"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gpumpi/host.h"

extern "C" const void* gpumpi_prog_lookup(const char* name);

// Rank 1 raises an MPI error; every other rank runs Allreduce, Bcast and
// Alltoall.  The launch must end with that error and no rank may read a
// collective table entry its peers never published.
GPUMPI_PROGRAM(sim_abort_collectives) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    int* buf = (int*)gpumpi_malloc(sizeof(int) * 4 * (uint64_t)P);
    for (int i = 0; i < 4 * P; i++) buf[i] = rank + i;
    if (rank == 1) {
        MPI_Abort(MPI_COMM_WORLD, 7);
    } else {
        for (int it = 0; it < 3; it++) {
            MPI_Allreduce(MPI_IN_PLACE, buf, P, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
            MPI_Bcast(buf, P, MPI_INT, 0, MPI_COMM_WORLD);
            MPI_Alltoall(buf, 1, MPI_INT, buf + 2 * P, 1, MPI_INT, MPI_COMM_WORLD);
        }
    }
    MPI_Finalize();
    return 0;
}

// Self-test of the happens-before checker: rank 0 writes a word of a
// shared buffer that rank 1 then reads, with (args[2] = 1) or without
// (args[2] = 0) an MPI_Barrier between the two; and rank 1 overwrites a word
// rank 0 read, with or without the barrier.  Only the variant without the
// barrier may be reported.
GPUMPI_PROGRAM(sim_race_selftest) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    volatile int* buf = prog_ptr<int>(1);
    const bool ordered = prog_int(2) != 0;
    if (rank == 0) {
        buf[0] = 42;
        buf[8] += 1;
    }
    if (ordered) MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 1) {
        buf[4] = buf[0];
        buf[8] = 7;
    }
    MPI_Finalize();
    return 0;
}

static int launch(const char* prog, int P, int threads, std::vector<uint64_t>& words, int* mpi_error = nullptr) {
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, P);
    cfg.threads_per_rank = threads;
    cfg.timeout_s = 0;
    gpumpi_result res;
    char name[64];
    std::snprintf(name, sizeof name, "%s", prog);
    char* av[] = {name, nullptr};
    const int rc = gpumpi_run(&cfg, gpumpi_prog_lookup(prog), 1, av, words.data(), &res);
    std::printf("%s P=%d rc=%d status=%d mpi_error=%d rank=%d line=%d msg='%s'\n", prog, P, rc, res.status,
                res.mpi_error, res.mpi_error_rank, res.mpi_error_line, res.mpi_error_msg);
    if (rc) std::printf("error: %s\n", gpumpi_last_error());
    std::fflush(stdout);
    if (mpi_error) *mpi_error = res.mpi_error;
    return rc;
}

static int alltoallv(const char* file, int threads) {
    FILE* f = std::fopen(file, "r");
    if (!f) return 2;
    int P = 0;
    if (std::fscanf(f, "%d", &P) != 1) return 2;
    std::vector<int> counts((size_t)P * P);
    for (int& c : counts) if (std::fscanf(f, "%d", &c) != 1) return 2;
    std::fclose(f);
    std::vector<long> srow(P, 0), rcol(P, 0);
    for (int s = 0; s < P; s++)
        for (int d = 0; d < P; d++) {
            srow[s] += counts[(size_t)s * P + d];
            rcol[d] += counts[(size_t)s * P + d];
        }
    long smax = 0, rmax = 0;
    for (int r = 0; r < P; r++) {
        smax = std::max(smax, srow[r]);
        rmax = std::max(rmax, rcol[r]);
    }
    const uint64_t sstride = ((uint64_t)smax * 4 + 255) / 256 * 256 + 256;
    const uint64_t rstride = ((uint64_t)rmax * 4 + 255) / 256 * 256 + 256;
    uint8_t *sbp = nullptr, *rbp = nullptr;
    if (gpumpi_device_alloc((void**)&sbp, sstride * P) || gpumpi_device_alloc((void**)&rbp, rstride * P)) return 4;
    std::memset(rbp, 0, rstride * P);
    struct View { uint8_t* p; uint8_t* data() { return p; } } sb{sbp}, rb{rbp};
    for (int r = 0; r < P; r++)
        for (long i = 0; i < srow[r]; i++) ((int32_t*)(sb.data() + r * sstride))[i] = (int32_t)(r * 1000003 + i * 7919);
    std::vector<uint64_t> words(64, 0);
    words[1] = 2;  // MPI_INT (datatype enum order)
    words[2] = (uint64_t)(uintptr_t)counts.data();
    words[4] = (uint64_t)(uintptr_t)sb.data();
    words[5] = (uint64_t)(uintptr_t)rb.data();
    words[6] = sstride;
    words[7] = rstride;
    words[10] = 1;
    if (launch("coll_alltoallv", P, threads, words)) return 1;
    long bad = 0;
    for (int r = 0; r < P; r++) {
        long o = 0;
        for (int s = 0; s < P; s++) {
            long sd = 0;
            for (int k = 0; k < r; k++) sd += counts[(size_t)s * P + k];
            for (int i = 0; i < counts[(size_t)s * P + r]; i++, o++)
                bad += ((int32_t*)(rb.data() + r * rstride))[o] != (int32_t)(s * 1000003 + (sd + i) * 7919);
        }
    }
    std::printf("alltoallv mismatches: %ld\n", bad);
    return bad ? 1 : 0;
}

static int blocks(int kind, int threads) {
    const int P = 6, blk = 5, root = 2;
    const uint64_t stride = 4 * 1024;
    std::vector<uint8_t> sb(stride * P), rb(stride * P, 0);
    for (int r = 0; r < P; r++)
        for (int i = 0; i < 1024; i++) ((int32_t*)(sb.data() + r * stride))[i] = i + 1000 * r;
    if (kind == 7 || kind == 8) rb = sb;  // in place: the receive buffer holds the blocks to send
    std::vector<uint64_t> words(64, 0);
    words[1] = (uint64_t)kind;
    words[2] = blk;
    words[3] = root;
    words[4] = (uint64_t)(uintptr_t)sb.data();
    words[5] = (uint64_t)(uintptr_t)rb.data();
    words[6] = stride;
    words[7] = stride;
    if (launch("coll_blocks", P, threads, words)) return 1;
    auto got = [&](int r, int i) { return ((int32_t*)(rb.data() + r * stride))[i]; };
    auto val = [&](int r, int i) { return i + 1000 * r; };
    long bad = 0;
    std::vector<int> cnt(P), dsp(P);
    for (int r = 0; r < P; r++) {
        cnt[r] = blk + r;
        dsp[r] = r ? dsp[r - 1] + cnt[r - 1] : 0;
    }
    switch (kind) {
    case 0: for (int r = 0; r < P; r++) for (int i = 0; i < blk; i++) bad += got(root, r * blk + i) != val(r, i); break;
    case 1: for (int r = 0; r < P; r++) for (int i = 0; i < cnt[r]; i++) bad += got(root, dsp[r] + i) != val(r, i); break;
    case 2: for (int r = 0; r < P; r++) for (int i = 0; i < blk; i++) bad += got(r, i) != val(root, r * blk + i); break;
    case 3: for (int r = 0; r < P; r++) for (int i = 0; i < cnt[r]; i++) bad += got(r, i) != val(root, dsp[r] + i); break;
    case 4: for (int d = 0; d < P; d++) for (int r = 0; r < P; r++) for (int i = 0; i < blk; i++) bad += got(d, r * blk + i) != val(r, i); break;
    case 5: for (int d = 0; d < P; d++) for (int r = 0; r < P; r++) for (int i = 0; i < cnt[r]; i++) bad += got(d, dsp[r] + i) != val(r, i); break;
    case 8:  // Alltoallv in place: ranks r and k exchange blk + (r + k) % 3 elements both ways
        for (int r = 0; r < P; r++) {
            int o = 0;
            for (int src = 0; src < P; src++) {
                int d = 0;
                for (int k = 0; k < r; k++) d += blk + (src + k) % 3;
                for (int i = 0; i < blk + (src + r) % 3; i++, o++) bad += got(r, o) != val(src, d + i);
            }
        }
        break;
    default: for (int d = 0; d < P; d++) for (int s = 0; s < P; s++) for (int i = 0; i < blk; i++) bad += got(d, s * blk + i) != val(s, d * blk + i); break;
    }
    std::printf("blocks kind %d mismatches: %ld\n", kind, bad);
    return bad ? 1 : 0;
}

// Compressed-wire Allreduce (coll_reduce with args[13] = wire) of P rows of n
// floats: every rank's output must be the same and within the wire's error
// of the exact sum.
static int wire(int wire_fmt, int P, int n, int threads) {
    const uint64_t stride = ((uint64_t)n * 4 + 255) / 256 * 256;
    uint8_t *ib = nullptr, *ob = nullptr;
    if (gpumpi_device_alloc((void**)&ib, stride * P) || gpumpi_device_alloc((void**)&ob, stride * P)) return 4;
    std::memset(ob, 0, stride * P);
    std::vector<double> exact(n, 0.0);
    for (int r = 0; r < P; r++)
        for (int i = 0; i < n; i++) {
            const float v = (float)((((int64_t)r * 7919 + (int64_t)i * 104729) % 2001) - 1000) / 37.0f;
            ((float*)(ib + r * stride))[i] = v;
            exact[i] += v;
        }
    std::vector<uint64_t> words(64, 0);
    words[1] = 12;  // MPI_FLOAT (datatype enum order)
    words[2] = 2;   // MPI_SUM (operator index)
    words[3] = (uint64_t)n;
    words[4] = (uint64_t)(uintptr_t)ib;
    words[5] = (uint64_t)(uintptr_t)ob;
    words[6] = stride;
    words[7] = (uint64_t)(int64_t)-1;  // no root: Allreduce
    words[9] = 1;
    words[10] = 1;
    words[13] = (uint64_t)wire_fmt;
    if (launch("coll_reduce", P, threads, words)) return 1;
    long bad = 0;
    double e2 = 0, x2 = 0;
    for (int r = 0; r < P; r++)
        for (int i = 0; i < n; i++) {
            const float g = ((float*)(ob + r * stride))[i];
            bad += std::memcmp(&g, ob + i * 4, 4) != 0;
            if (r == 0) {
                e2 += (g - exact[i]) * (g - exact[i]);
                x2 += exact[i] * exact[i];
            }
        }
    const double rel = x2 > 0 ? std::sqrt(e2 / x2) : 0;
    std::printf("wire %d P=%d n=%d: rank mismatches %ld, relative error %.3g\n", wire_fmt, P, n, bad, rel);
    return bad == 0 && rel < (wire_fmt == 1 ? 1e-2 : 6e-2) ? 0 : 1;
}

static int race_selftest(int ordered) {
    int* buf = nullptr;
    if (gpumpi_device_alloc((void**)&buf, 256)) return 4;
    std::memset(buf, 0, 256);
    std::vector<uint64_t> words(64, 0);
    words[1] = (uint64_t)(uintptr_t)buf;
    words[2] = (uint64_t)ordered;
    return launch("sim_race_selftest", 2, 64, words);
}

// coll_blocks kind 9: an Alltoall whose send blocks exceed the receive
// blocks must end the launch with MPI_ERR_TRUNCATE.
static int truncate_case(int threads) {
    const int P = 6, blk = 5;
    const uint64_t stride = 4 * 1024;
    std::vector<uint8_t> sb(stride * P, 0), rb(stride * P, 0);
    std::vector<uint64_t> words(64, 0);
    words[1] = 9;
    words[2] = blk;
    words[4] = (uint64_t)(uintptr_t)sb.data();
    words[5] = (uint64_t)(uintptr_t)rb.data();
    words[6] = stride;
    words[7] = stride;
    int err = 0;
    const int rc = launch("coll_blocks", P, threads, words, &err);
    std::printf("truncate case: rc=%d mpi_error=%d (expect rc=%d, mpi_error=%d)\n", rc, err, GPUMPI_EDEVICE,
                MPI_ERR_TRUNCATE);
    return (rc == GPUMPI_EDEVICE && err == MPI_ERR_TRUNCATE) ? 0 : 1;
}

static int abort_case(int P, int threads) {
    std::vector<uint64_t> words(64, 0);
    int err = 0;
    const int rc = launch("sim_abort_collectives", P, threads, words, &err);
    std::printf("abort case: rc=%d mpi_error=%d (expect rc=%d, mpi_error=7)\n", rc, err, GPUMPI_EDEVICE);
    return (rc == GPUMPI_EDEVICE && err == 7) ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc >= 4 && !std::strcmp(argv[1], "abort")) return abort_case(std::atoi(argv[2]), std::atoi(argv[3]));
    if (argc >= 4 && !std::strcmp(argv[1], "alltoallv")) return alltoallv(argv[2], std::atoi(argv[3]));
    if (argc >= 4 && !std::strcmp(argv[1], "blocks")) return blocks(std::atoi(argv[2]), std::atoi(argv[3]));
    if (argc >= 3 && !std::strcmp(argv[1], "race_selftest")) return race_selftest(std::atoi(argv[2]));
    if (argc >= 3 && !std::strcmp(argv[1], "truncate")) return truncate_case(std::atoi(argv[2]));
    if (argc >= 6 && !std::strcmp(argv[1], "wire"))
        return wire(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    std::fprintf(stderr, "usage: run_coll alltoallv COUNTS THREADS | blocks KIND THREADS | abort P THREADS\n");
    return 2;
}
