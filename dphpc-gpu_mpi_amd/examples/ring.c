/* A C MPI program for the source converter (dphpc-gpu_mpi_amd/converter.py)
 * and its tests, built by gpumpicc into lib/ring: globals with and
 * without initialisers, a static local, a const global, a function
 * prototype, an implicit void* conversion from malloc, a variable named with
 * a C++ keyword, argv parsing with atoi, and formatted output.  Each rank
 * sends a token around the ring, then all reduce a checksum. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mpi.h"

static int rounds = 3;
int hops;
long checksum[2] = {7, 11};
const int SCALE = 10;

static int next_of(int r, int n);

static int next_of(int r, int n) { return (r + 1) % n; }

static int count_calls(void) {
    static int calls = 0;
    return ++calls;
}

int main(int argc, char* argv[]) {
    int rank, size, i;
    int class = 0;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc > 1) rounds = atoi(argv[1]);
    int* buf = malloc(sizeof(int) * 4);
    memset(buf, 0, sizeof(int) * 4);
    /* heap message buffers: a receive into a stack variable is matched
     * later than the call that posted it, which the CPU simulation of the
     * device cannot model (each simulated lane has its own stack) */
    int* msg = calloc(2, sizeof(int));
    for (i = 0; i < rounds; i++) {
        msg[0] = rank * SCALE + i;
        msg[1] = -1;
        MPI_Sendrecv(&msg[0], 1, MPI_INT, next_of(rank, size), i, &msg[1], 1, MPI_INT, (rank + size - 1) % size, i,
                     MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        buf[i % 4] += msg[1];
        hops++;
        class += count_calls();
    }
    checksum[0] += buf[0] + buf[1] + buf[2] + buf[3] + hops + class;
    long total = 0;
    MPI_Allreduce(&checksum[0], &total, 1, MPI_LONG, MPI_SUM, MPI_COMM_WORLD);
    char line[64];
    sprintf(line, "%ld/%d", checksum[0], hops);
    if (rank == 0) printf("ring: size %d rounds %d total %ld first %s\n", size, rounds, total, line);
    fflush(stdout);
    free(msg);
    free(buf);
    MPI_Finalize();
    return 0;
}
