/*
 * Host-MPI golden records of the point-to-point / request / communicator
 * semantics scenarios (test infrastructure).
 *
 * The scenarios are the very text the device rank program `semantics` runs
 * (dphpc-gpu_mpi_amd/csrc/progs/semantics_scenarios.h); here they run under
 * the host MPI library, and rank 0 prints every rank's record as one JSON
 * object:  {"P": P, "mask": M, "records": [[...rank 0...], ...]}.
 *
 * usage: mpirun -np P semantics_golden [mask]
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>

#define SEM_FN static
#define SEM_ALLOC(b) malloc(b)
#define SEM_FREE(p, b) free(p)
#include "../../dphpc-gpu_mpi_amd/csrc/progs/semantics_scenarios.h"

#define CAP 512

int main(int argc, char** argv) {
    int rank, P, i, q;
    int mask = argc > 1 ? atoi(argv[1]) : (1 << SEM_NUM_SCENARIOS) - 1;
    struct sem_rec r;
    int *all = NULL, *counts = NULL;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    r.v = (int*)calloc(CAP, sizeof(int));
    r.n = 0;
    r.cap = CAP;
    sem_run(&r, rank, P, mask);
    if (rank == 0) {
        all = (int*)malloc(sizeof(int) * CAP * (size_t)P);
        counts = (int*)malloc(sizeof(int) * (size_t)P);
    }
    MPI_Gather(&r.n, 1, MPI_INT, counts, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Gather(r.v, CAP, MPI_INT, all, CAP, MPI_INT, 0, MPI_COMM_WORLD);
    if (rank == 0) {
        printf("{\"P\": %d, \"mask\": %d, \"records\": [", P, mask);
        for (q = 0; q < P; q++) {
            const int n = counts[q] < CAP ? counts[q] : CAP;
            printf("%s[", q ? ", " : "");
            for (i = 0; i < n; i++) printf("%s%d", i ? ", " : "", all[(size_t)q * CAP + i]);
            printf("]");
        }
        printf("]}\n");
        free(all);
        free(counts);
    }
    free(r.v);
    MPI_Finalize();
    return 0;
}
