/* Contiguous derived datatypes through a C MPI program: sizes, nesting,
 * point to point with a stack buffer, Bcast, Allreduce and Reduce with a
 * user-defined operator (the predefined ones are for basic datatypes),
 * MPI_Get_count in derived units, MPI_Type_free.  Prints
 * "derived_types: ok" on rank 0 when every check holds (under host MPI
 * too), else the failed check, and returns non-zero. */
#include <stdio.h>
#include "mpi.h"

/* inout += in over every int of len elements of the derived type */
static void isum(void* in, void* inout, int* len, MPI_Datatype* dt) {
    int sz = 0, k;
    MPI_Type_size(*dt, &sz);
    for (k = 0; k < *len * sz / (int)sizeof(int); k++) ((int*)inout)[k] += ((int*)in)[k];
}

static void imax(void* in, void* inout, int* len, MPI_Datatype* dt) {
    int sz = 0, k;
    MPI_Type_size(*dt, &sz);
    for (k = 0; k < *len * sz / (int)sizeof(int); k++)
        if (((int*)in)[k] > ((int*)inout)[k]) ((int*)inout)[k] = ((int*)in)[k];
}

#define CHECK(x) do { if (!(x)) { printf("derived_types: rank %d failed %s\n", rank, #x); bad = 1; } } while (0)

int main(int argc, char** argv) {
    int rank, size, bad = 0, s = 0, n = -1, i;
    MPI_Datatype tri, six;
    MPI_Op sum, max;
    MPI_Status st;
    int buf[12], out[12];
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    MPI_Type_contiguous(3, MPI_INT, &tri);
    MPI_Type_commit(&tri);
    MPI_Type_contiguous(2, tri, &six);
    MPI_Type_commit(&six);
    MPI_Type_size(tri, &s);
    CHECK(s == 12);
    MPI_Type_size(six, &s);
    CHECK(s == 24);
    for (i = 0; i < 12; i++) buf[i] = rank * 100 + i;
    if (rank == 0) {
        MPI_Send(buf, 2, tri, 1 % size, 7, MPI_COMM_WORLD);
    }
    if (rank == 1 % size) {
        int got[6];
        MPI_Recv(got, 1, six, 0, 7, MPI_COMM_WORLD, &st);
        MPI_Get_count(&st, tri, &n);
        CHECK(n == 2);
        for (i = 0; i < 6; i++) CHECK(got[i] == i);
    }
    if (rank == 0)
        for (i = 0; i < 12; i++) buf[i] = 1000 + i;
    MPI_Bcast(buf, 2, six, 0, MPI_COMM_WORLD);
    for (i = 0; i < 12; i++) CHECK(buf[i] == 1000 + i);
    for (i = 0; i < 12; i++) buf[i] = rank + i;
    MPI_Op_create(isum, 1, &sum);
    MPI_Op_create(imax, 1, &max);
    MPI_Allreduce(buf, out, 4, tri, sum, MPI_COMM_WORLD);
    for (i = 0; i < 12; i++) CHECK(out[i] == size * (size - 1) / 2 + size * i);
    for (i = 0; i < 12; i++) out[i] = -1;
    MPI_Reduce(buf, out, 2, six, max, size - 1, MPI_COMM_WORLD);
    if (rank == size - 1)
        for (i = 0; i < 12; i++) CHECK(out[i] == size - 1 + i);
    MPI_Type_free(&six);
    CHECK(six == MPI_DATATYPE_NULL);
    MPI_Type_free(&tri);
    MPI_Op_free(&sum);
    MPI_Op_free(&max);
    MPI_Finalize();
    if (rank == 0 && !bad) printf("derived_types: ok\n");
    return bad;
}
