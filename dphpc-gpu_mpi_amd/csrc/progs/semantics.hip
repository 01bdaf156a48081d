// The point-to-point / request / communicator semantics scenarios
// (semantics_scenarios.h) as a rank program of this runtime; the same text
// under host MPI writes the golden records the GPU tests compare with.
#include "registry.hpp"

#define SEM_FN __device__ inline
#define SEM_ALLOC(b) gpumpi_malloc(b)
#define SEM_FREE(p, b) gpumpi_free((p), (b))
#include "semantics_scenarios.h"

// args: [0] records (int[P * cap]), [1] cap, [2] scenario mask,
// [3] record lengths (int[P])
GPUMPI_PROGRAM(semantics) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int cap = (int)prog_int(1);
    struct sem_rec r;
    r.v = prog_ptr<int>(0) + (uint64_t)rank * (uint64_t)cap;
    r.n = 0;
    r.cap = cap;
    sem_run(&r, rank, P, (int)prog_int(2));
    prog_ptr<int>(3)[rank] = r.n;
    MPI_Finalize();
    return 0;
}

// A receive whose buffer is smaller than the message (MPI_ERR_TRUNCATE): the
// launch must end with the error record naming the receiving rank.
// args: [1] 0 = eager size, 1 = rendezvous size
GPUMPI_PROGRAM(sem_truncate) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    const int n = prog_int(1) ? 4096 : 10;
    int* buf = (int*)gpumpi_malloc(sizeof(int) * (uint64_t)n);
    for (int i = 0; i < n; i++) buf[i] = i;
    if (rank == 1) MPI_Send(buf, n, MPI_INT, 0, 1, MPI_COMM_WORLD);
    if (rank == 0) MPI_Recv(buf, n / 2, MPI_INT, 1, 1, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
    MPI_Finalize();
    return 0;
}
