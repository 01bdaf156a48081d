/*
 * Host-MPI point-to-point harness (test infrastructure / CPU baseline).
 *
 * The same workload the device runs, under the host MPI library:
 *   ring:     every rank Irecv(left) + Isend(right) + Waitall, `iters` times
 *   pingpong: pairs (2p, 2p+1) Send/Recv | Recv/Send, `iters` round trips
 * Timing follows benchmarks/mpi_communication/mpi_communication.c:16-51
 * (MPI_Wtime around the loop, max over ranks).  Payloads are the oracle's
 * seeded words (oracle.c orc_payload_word) and every received buffer is
 * checked against the sender's payload.
 *
 * usage: p2p ring|pingpong BYTES ITERS WARMUP SEED
 * prints one JSON line from rank 0.
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static uint64_t word(uint64_t seed, uint64_t rank, uint64_t i) {
    return mix64(seed + rank * 0xD1B54A32D192ED03ULL + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
static void fill(uint8_t* p, size_t n, uint64_t seed, int rank) {
    size_t nw = n / 8;
    for (size_t i = 0; i < nw; i++) {
        uint64_t w = word(seed, (uint64_t)rank, i);
        memcpy(p + 8 * i, &w, 8);
    }
    if (n % 8) {
        uint64_t w = word(seed, (uint64_t)rank, nw);
        memcpy(p + 8 * nw, &w, n % 8);
    }
}
static long long check(const uint8_t* p, size_t n, uint64_t seed, int rank) {
    uint8_t* e = malloc(n ? n : 1);
    fill(e, n, seed, rank);
    long long bad = 0;
    for (size_t i = 0; i < n; i++) bad += p[i] != e[i];
    free(e);
    return bad;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 6) {
        if (rank == 0) fprintf(stderr, "usage: %s ring|pingpong BYTES ITERS WARMUP SEED\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    const int ring = strcmp(argv[1], "ring") == 0;
    const size_t n = (size_t)atoll(argv[2]);
    const int iters = atoi(argv[3]), warm = atoi(argv[4]);
    const uint64_t seed = (uint64_t)strtoull(argv[5], NULL, 0);
    uint8_t* sb = malloc(n ? n : 1);
    uint8_t* rb = calloc(n ? n : 1, 1);
    fill(sb, n, seed, rank);
    double t0 = 0, t = 0;
    long long bad = 0;
    if (ring) {
        const int right = (rank + 1) % size, left = (rank + size - 1) % size;
        for (int it = 0; it < warm + iters; it++) {
            if (it == warm) {
                MPI_Barrier(MPI_COMM_WORLD);
                t0 = MPI_Wtime();
            }
            MPI_Request rq[2];
            MPI_Irecv(rb, (int)n, MPI_BYTE, left, 7, MPI_COMM_WORLD, &rq[0]);
            MPI_Isend(sb, (int)n, MPI_BYTE, right, 7, MPI_COMM_WORLD, &rq[1]);
            MPI_Waitall(2, rq, MPI_STATUSES_IGNORE);
        }
        t = MPI_Wtime() - t0;
        bad = check(rb, n, seed, left);
    } else {
        const int peer = rank ^ 1;
        const int active = peer < size;
        MPI_Barrier(MPI_COMM_WORLD);
        for (int it = 0; active && it < warm + iters; it++) {
            if (it == warm) t0 = MPI_Wtime();
            if ((rank & 1) == 0) {
                MPI_Send(sb, (int)n, MPI_BYTE, peer, 10, MPI_COMM_WORLD);
                MPI_Recv(rb, (int)n, MPI_BYTE, peer, 10, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
            } else {
                MPI_Recv(rb, (int)n, MPI_BYTE, peer, 10, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                MPI_Send(sb, (int)n, MPI_BYTE, peer, 10, MPI_COMM_WORLD);
            }
        }
        t = active ? MPI_Wtime() - t0 : 0.0;
        if (active) bad = check(rb, n, seed, peer);
    }
    double tmax = 0;
    long long bad_all = 0;
    MPI_Reduce(&t, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    MPI_Reduce(&bad, &bad_all, 1, MPI_LONG_LONG, MPI_SUM, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        double bytes = ring ? 2.0 * size * (double)n * iters : 2.0 * (double)n * iters;
        printf("{\"mode\": \"%s\", \"ranks\": %d, \"bytes\": %zu, \"iters\": %d, \"seconds\": %.6f, "
               "\"gbps\": %.4f, \"us_per_iter\": %.4f, \"mismatches\": %lld}\n",
               argv[1], size, n, iters, tmax, bytes / tmax / 1e9, tmax / iters * 1e6, bad_all);
    }
    free(sb);
    free(rb);
    MPI_Finalize();
    return 0;
}
