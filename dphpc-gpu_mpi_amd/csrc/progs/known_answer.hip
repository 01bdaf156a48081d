// Rank programs restating the reference's GPU known-answer tests on the
// device MPI API.  Each writes ok[world rank] = 1 on success (args[0] = ok).
//   tests/cuda_mpi_test.cu:5-317, tests/collectives_test.cu:6-288,
//   tests/communicator_test.cu:5-152.
#include "registry.hpp"

__device__ inline void set_ok(bool ok) {
    int* okv = prog_ptr<int>(0);
    if (gpumpi_is_leader_lane()) okv[gpumpi_world_rank()] = ok ? 1 : 0;
}

// cuda_mpi_test.cu:5-36: one int (3456) from rank 0 to rank 1.
GPUMPI_PROGRAM(ka_single_int) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank == 0) {
        int x = 3456;
        MPI_Send(&x, 1, MPI_INT, 1, 15, MPI_COMM_WORLD);
    } else if (rank == 1) {
        int x = 0;
        MPI_Recv(&x, 1, MPI_INT, 0, 15, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        ok = x == 3456;
    }
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// cuda_mpi_test.cu:38-73: int[16] = 0..15 from a stack array.
GPUMPI_PROGRAM(ka_array16) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank == 0) {
        int x[16];
        for (int i = 0; i < 16; i++) x[i] = i;
        MPI_Send(x, 16, MPI_INT, 1, 15, MPI_COMM_WORLD);
    } else if (rank == 1) {
        int x[16] = {};
        MPI_Recv(x, 16, MPI_INT, 0, 15, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        for (int i = 0; i < 16; i++) ok = ok && x[i] == i;
    }
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// cuda_mpi_test.cu:75-127: both ranks Irecv then Isend (+i / -i).
GPUMPI_PROGRAM(ka_sendrecv_pm) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank < 2) {
        int other = 1 - rank;
        int x[16], y[16] = {};
        for (int i = 0; i < 16; i++) x[i] = rank == 0 ? i : -i;
        MPI_Request rq[2];
        MPI_Irecv(y, 16, MPI_INT, other, 15, MPI_COMM_WORLD, &rq[0]);
        MPI_Isend(x, 16, MPI_INT, other, 15, MPI_COMM_WORLD, &rq[1]);
        MPI_Wait(&rq[0], MPI_STATUS_IGNORE);
        for (int i = 0; i < 16; i++) ok = ok && y[i] == (rank == 0 ? -i : i);
        MPI_Wait(&rq[1], MPI_STATUS_IGNORE);
    }
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// cuda_mpi_test.cu:129-174: 5 tags with growing sizes, waits out of order.
GPUMPI_PROGRAM(ka_repeat_tags) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    const int other = (rank + 1) % 2;
    const int reps = 5, n = 1 << (reps - 1);
    int local[n], remote[n];
    for (int i = 0; i < n; i++) {
        local[i] = (rank + 1) * i;
        remote[i] = 0;
    }
    MPI_Request s[reps], r[reps];
    for (int i = 0; i < reps; i++) {
        MPI_Isend(local, 1 << i, MPI_INT, other, i + 10, MPI_COMM_WORLD, &s[i]);
        MPI_Irecv(remote, 1 << i, MPI_INT, other, i + 10, MPI_COMM_WORLD, &r[i]);
    }
    for (int i = 0; i < reps; i++) {
        MPI_Wait(&s[i], MPI_STATUS_IGNORE);
        MPI_Wait(&r[i], MPI_STATUS_IGNORE);
    }
    bool ok = true;
    for (int i = 0; i < n; i++) ok = ok && remote[i] == (other + 1) * i;
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// cuda_mpi_test.cu:177-233: 3 ranks, every rank talks to both neighbours.
GPUMPI_PROGRAM(ka_network_flood) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int next = (rank + 1) % size, prev = (rank + size - 1) % size;
    const int reps = 5, n = 1 << (reps - 1);
    int local[n], pd[n], nd[n];
    for (int i = 0; i < n; i++) {
        local[i] = (rank + 1) * i;
        pd[i] = nd[i] = 0;
    }
    MPI_Request sn[reps], rn[reps], sp[reps], rp[reps];
    for (int i = 0; i < reps; i++) {
        MPI_Isend(local, 1 << i, MPI_INT, next, i + 10, MPI_COMM_WORLD, &sn[i]);
        MPI_Irecv(nd, 1 << i, MPI_INT, next, i + 10, MPI_COMM_WORLD, &rn[i]);
        MPI_Isend(local, 1 << i, MPI_INT, prev, i + 10, MPI_COMM_WORLD, &sp[i]);
        MPI_Irecv(pd, 1 << i, MPI_INT, prev, i + 10, MPI_COMM_WORLD, &rp[i]);
    }
    for (int i = 0; i < reps; i++) {
        MPI_Wait(&rp[i], MPI_STATUS_IGNORE);
        MPI_Wait(&sn[i], MPI_STATUS_IGNORE);
        MPI_Wait(&sp[i], MPI_STATUS_IGNORE);
        MPI_Wait(&rn[i], MPI_STATUS_IGNORE);
    }
    bool ok = true;
    for (int i = 0; i < n; i++) ok = ok && pd[i] == (prev + 1) * i && nd[i] == (next + 1) * i;
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// cuda_mpi_test.cu:235-317: every pair, 20 repeats with the same tag; the
// repeats must arrive in order (non-overtaking).  Buffers on the rank heap.
GPUMPI_PROGRAM(ka_all_to_all_fifo) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int reps = 20, n = 16;
    int* local = (int*)gpumpi_malloc(sizeof(int) * n * reps * size);
    int* remote = (int*)gpumpi_malloc(sizeof(int) * n * reps * size);
    MPI_Request* sr = (MPI_Request*)gpumpi_malloc(sizeof(MPI_Request) * reps * size);
    MPI_Request* rr = (MPI_Request*)gpumpi_malloc(sizeof(MPI_Request) * reps * size);
    for (int r = 0; r < size; r++)
        for (int k = 0; k < reps; k++)
            for (int i = 0; i < n; i++) {
                const int idx = i + k * n + r * n * reps;
                local[idx] = i * k * r;
                remote[idx] = -1;
            }
    for (int k = 0; k < reps; k++)
        for (int r = 0; r < size; r++) {
            sr[r + k * size] = MPI_REQUEST_NULL;
            rr[r + k * size] = MPI_REQUEST_NULL;
            if (r == rank) continue;
            const int idx = k * n + r * n * reps;
            MPI_Isend(local + idx, n, MPI_INT, r, 15, MPI_COMM_WORLD, &sr[r + k * size]);
            MPI_Irecv(remote + idx, n, MPI_INT, r, 15, MPI_COMM_WORLD, &rr[r + k * size]);
        }
    MPI_Waitall(reps * size, sr, MPI_STATUSES_IGNORE);
    MPI_Waitall(reps * size, rr, MPI_STATUSES_IGNORE);
    bool ok = true;
    int64_t* diag = prog_ptr<int64_t>(1);  // optional: [rank*64]: count, then (k, r, i, got) records
    int nbad = 0;
    for (int k = 0; k < reps; k++)
        for (int r = 0; r < size; r++) {
            if (r == rank) continue;
            for (int i = 0; i < n; i++) {
                const int got = remote[i + k * n + r * n * reps];
                if (got != i * k * rank) {
                    ok = false;
                    if (diag && nbad < 15) {
                        int64_t* d = diag + (int64_t)rank * 64 + 1 + 4 * nbad;
                        d[0] = k;
                        d[1] = r;
                        d[2] = i;
                        d[3] = got;
                    }
                    nbad++;
                }
            }
        }
    if (diag) diag[(int64_t)rank * 64] = nbad;
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// collectives_test.cu:6-54: Reduce PROD of int16 with wrap-around.
GPUMPI_PROGRAM(ka_reduce_prod_i16) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 14;
    int16_t buf[n], out[n] = {};
    for (int i = 0; i < n; i++) buf[i] = (int16_t)((rank + 1) * (i + 1));
    MPI_Reduce(buf, out, n, MPI_INT16_T, MPI_PROD, 0, MPI_COMM_WORLD);
    bool ok = true;
    if (rank == 0) {
        for (int i = 0; i < n; i++) {
            int16_t e = 1;
            for (int r = 0; r < size; r++) e = (int16_t)(e * (int16_t)((r + 1) * (i + 1)));
            ok = ok && out[i] == e;
        }
    }
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// collectives_test.cu:57-98: Allreduce SUM of doubles.
GPUMPI_PROGRAM(ka_allreduce_double) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 17;
    double buf[n], out[n] = {};
    for (int i = 0; i < n; i++) buf[i] = (double)((rank + 1) * (i + 1));
    MPI_Allreduce(buf, out, n, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < n; i++) {
        double e = 0;
        for (int r = 0; r < size; r++) e += (double)((r + 1) * (i + 1));
        ok = ok && out[i] == e;
    }
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// collectives_test.cu:100-141: Allgather of int[7].
GPUMPI_PROGRAM(ka_allgather) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 7;
    int* buf = (int*)gpumpi_malloc(sizeof(int) * n);
    int* out = (int*)gpumpi_malloc(sizeof(int) * n * size);
    for (int i = 0; i < n; i++) buf[i] = (rank + 1) * (i + 1);
    MPI_Allgather(buf, n, MPI_INT, out, n, MPI_INT, MPI_COMM_WORLD);
    bool ok = true;
    for (int r = 0; r < size; r++)
        for (int i = 0; i < n; i++) ok = ok && out[r * n + i] == (r + 1) * (i + 1);
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// collectives_test.cu:144-196: Scatter from root 2.
GPUMPI_PROGRAM(ka_scatter) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int root = 2, n = 4;
    int* sb = (int*)gpumpi_malloc(sizeof(int) * n * size);
    int rb[n];
    for (int j = 0; j < size; j++)
        for (int i = 0; i < n; i++) sb[i + j * n] = rank == root ? j : -2;
    for (int i = 0; i < n; i++) rb[i] = -1;
    MPI_Scatter(sb, n, MPI_INT, rb, n, MPI_INT, root, MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < n; i++) ok = ok && rb[i] == rank;
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// collectives_test.cu:199-288: Scatterv with sizes i+1.
GPUMPI_PROGRAM(ka_scatterv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int root = 2;
    int* sizes = (int*)gpumpi_malloc(sizeof(int) * size);
    int* displs = (int*)gpumpi_malloc(sizeof(int) * size);
    for (int i = 0; i < size; i++) {
        sizes[i] = i + 1;
        displs[i] = i == 0 ? 0 : displs[i - 1] + sizes[i - 1];
    }
    const int total = displs[size - 1] + sizes[size - 1];
    int* sb = (int*)gpumpi_malloc(sizeof(int) * total);
    int* rb = (int*)gpumpi_malloc(sizeof(int) * sizes[rank]);
    for (int j = 0; j < size; j++)
        for (int i = 0; i < sizes[j]; i++) sb[displs[j] + i] = rank == root ? j : -2;
    for (int i = 0; i < sizes[rank]; i++) rb[i] = -1;
    MPI_Scatterv(sb, sizes, displs, MPI_INT, rb, sizes[rank], MPI_INT, root, MPI_COMM_WORLD);
    MPI_Barrier(MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < sizes[rank]; i++) ok = ok && rb[i] == rank;
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// communicator_test.cu:5-25: rank in WORLD equals the launch rank.
GPUMPI_PROGRAM(ka_comm_rank) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = rank == gpumpi_world_rank();
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// communicator_test.cu:27-90: same tag on three dup'd communicators stays
// isolated.
GPUMPI_PROGRAM(ka_comm_dup_isolation) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm c1, c2, c3;
    MPI_Comm_dup(MPI_COMM_WORLD, &c1);
    MPI_Comm_dup(MPI_COMM_WORLD, &c2);
    MPI_Comm_dup(c2, &c3);
    MPI_Request rq[3] = {MPI_REQUEST_NULL, MPI_REQUEST_NULL, MPI_REQUEST_NULL};
    int16_t data[30] = {};
    if (rank == 0) {
        for (int i = 0; i < 30; i++) data[i] = (int16_t)(i + 1);
        MPI_Isend(data + 0, 10, MPI_INT16_T, 1, 17, c1, &rq[0]);
        MPI_Irecv(data + 10, 10, MPI_INT16_T, 1, 17, c2, &rq[1]);
        MPI_Isend(data + 20, 10, MPI_INT16_T, 1, 17, c3, &rq[2]);
    } else if (rank == 1) {
        for (int i = 0; i < 30; i++) data[i] = (int16_t)-(i + 1);
        MPI_Irecv(data + 20, 10, MPI_INT16_T, 0, 17, c3, &rq[2]);
        MPI_Irecv(data + 0, 10, MPI_INT16_T, 0, 17, c1, &rq[0]);
        MPI_Isend(data + 10, 10, MPI_INT16_T, 0, 17, c2, &rq[1]);
    }
    MPI_Status st[3];
    MPI_Waitall(3, rq, st);
    bool ok = true;
    if (rank == 0) {
        for (int i = 10; i < 20; i++) ok = ok && data[i] == -(i + 1);
    } else if (rank == 1) {
        for (int i = 0; i < 10; i++) ok = ok && data[i] == i + 1;
        for (int i = 20; i < 30; i++) ok = ok && data[i] == i + 1;
    }
    MPI_Comm_free(&c1);
    MPI_Comm_free(&c2);
    MPI_Comm_free(&c3);
    MPI_Finalize();
    set_ok(ok);
    return 0;
}

// communicator_test.cu:93-152: Comm_split colour/key table on 10 ranks.
GPUMPI_PROGRAM(ka_comm_split) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    int color, key;
    if (rank < 3) {
        color = 0;
        key = rank;
    } else if (rank < 6) {
        color = 1;
        key = 6 - rank;
    } else if (rank < 9) {
        color = 2;
        key = 1;
    } else {
        color = 2;
        key = 0;
    }
    MPI_Comm comm;
    MPI_Comm_split(MPI_COMM_WORLD, color, key, &comm);
    int nr = -1, ns = -1;
    MPI_Comm_rank(comm, &nr);
    MPI_Comm_size(comm, &ns);
    bool ok;
    if (rank < 3) ok = nr == rank && ns == 3;
    else if (rank < 6) ok = nr == 5 - rank && ns == 3;
    else if (rank < 9) ok = nr == rank - 5 && ns == 4;
    else ok = nr == 0 && ns == 4;
    // The new communicator must carry traffic and collectives.
    int v = rank, sum = 0;
    MPI_Allreduce(&v, &sum, 1, MPI_INT, MPI_SUM, comm);
    int expect = 0;
    for (int r = 0; r < 10; r++) {
        int c = r < 3 ? 0 : r < 6 ? 1 : 2;
        if (c == color) expect += r;
    }
    ok = ok && sum == expect;
    MPI_Comm_free(&comm);
    MPI_Finalize();
    set_ok(ok);
    return 0;
}
