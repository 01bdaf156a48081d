/* A C MPI program whose ranks leave through exit(): rank 1 calls exit(3)
 * from a helper function after MPI_Finalize, the others return 0 from main.
 * Under the device runtime exit() ends the calling rank's program with that
 * value, so the launch reports one rank with a non-zero return. */
#include <stdio.h>
#include <stdlib.h>
#include "mpi.h"

static void leave(int code) {
    printf("rank leaving with %d\n", code);
    exit(code);
}

int main(int argc, char** argv) {
    int rank, size, sum = 0;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    MPI_Allreduce(&rank, &sum, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    if (rank == 0) printf("exit_early: size %d sum %d\n", size, sum);
    MPI_Finalize();
    if (rank == 1) leave(3);
    printf("never printed by rank 1: %d\n", rank);
    return 0;
}
