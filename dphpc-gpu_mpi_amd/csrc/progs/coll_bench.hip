// Collective workloads over host-provided buffers: used by the parity tests
// (results compared with the CPU oracle) and by the benchmarks.
// Buffers of rank r live at base + r * stride.
#include "registry.hpp"

// Communicator selection: 0 WORLD, 1 split by colour = rank % ncolors
// (key = rank), 2 a dup of WORLD.
__device__ inline MPI_Comm pick_comm(int mode, int ncolors) {
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm c = MPI_COMM_WORLD;
    if (mode == 1) MPI_Comm_split(MPI_COMM_WORLD, rank % ncolors, rank, &c);
    else if (mode == 2) MPI_Comm_dup(MPI_COMM_WORLD, &c);
    return c;
}

// args: [1] dtype, [2] op index (0..11), [3] count, [4] in base, [5] out
// base, [6] stride, [7] root (-1: Allreduce), [8] comm mode, [9] ncolors,
// [10] iters, [11] times (double[P]), [12] in place, [13] wire format
// (0 exact, 1 BF16, 2 FP8 e4m3: compressed Allreduce, MPI_FLOAT/SUM only)
GPUMPI_PROGRAM(coll_reduce) {
    MPI_Init(&argc, &argv);
    int wrank;
    MPI_Comm_rank(MPI_COMM_WORLD, &wrank);
    const MPI_Datatype dt = (MPI_Datatype)prog_int(1);
    const MPI_Op op = GPUMPI_OP_HANDLE(prog_int(2));
    const int count = (int)prog_int(3);
    const uint64_t stride = prog_args()[6];
    uint8_t* in = prog_ptr<uint8_t>(4) + (uint64_t)wrank * stride;
    uint8_t* out = prog_ptr<uint8_t>(5) + (uint64_t)wrank * stride;
    const int root = (int)prog_int(7);
    const int iters = (int)prog_int(10);
    double* times = prog_ptr<double>(11);
    const bool in_place = prog_int(12) != 0;
    const int wire = (int)prog_int(13);
    MPI_Comm comm = pick_comm((int)prog_int(8), (int)prog_int(9));
    int crank;
    MPI_Comm_rank(comm, &crank);
    MPI_Barrier(comm);
    const double t0 = MPI_Wtime();
    for (int it = 0; it < iters; it++) {
        const void* sb = in_place && (root < 0 || crank == root) ? MPI_IN_PLACE : in;
        if (root < 0 && wire) GPUMPI_Allreduce_wire(sb, out, count, dt, op, comm, wire);
        else if (root < 0) MPI_Allreduce(sb, out, count, dt, op, comm);
        else MPI_Reduce(sb, out, count, dt, op, root, comm);
    }
    const double t1 = MPI_Wtime();
    if (times && gpumpi_is_leader_lane()) times[wrank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

// args: [1] bytes, [4] buf base, [6] stride, [7] root, [8] comm mode,
// [9] ncolors, [10] iters, [11] times
GPUMPI_PROGRAM(coll_bcast) {
    MPI_Init(&argc, &argv);
    int wrank;
    MPI_Comm_rank(MPI_COMM_WORLD, &wrank);
    const int bytes = (int)prog_int(1);
    uint8_t* buf = prog_ptr<uint8_t>(4) + (uint64_t)wrank * prog_args()[6];
    const int root = (int)prog_int(7);
    const int iters = (int)prog_int(10);
    double* times = prog_ptr<double>(11);
    MPI_Comm comm = pick_comm((int)prog_int(8), (int)prog_int(9));
    MPI_Barrier(comm);
    const double t0 = MPI_Wtime();
    for (int it = 0; it < iters; it++) MPI_Bcast(buf, bytes, MPI_BYTE, root, comm);
    const double t1 = MPI_Wtime();
    if (times && gpumpi_is_leader_lane()) times[wrank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

// Alltoallv with a host-chosen count matrix counts[src * P + dst] (elements
// of dtype).  args: [1] dtype, [2] counts ptr (int[P*P]), [4] send base,
// [5] recv base, [6] send stride, [7] recv stride, [10] iters, [11] times
GPUMPI_PROGRAM(coll_alltoallv) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const MPI_Datatype dt = (MPI_Datatype)prog_int(1);
    const int* counts = prog_ptr<int>(2);
    uint8_t* sb = prog_ptr<uint8_t>(4) + (uint64_t)rank * prog_args()[6];
    uint8_t* rb = prog_ptr<uint8_t>(5) + (uint64_t)rank * prog_args()[7];
    const int iters = (int)prog_int(10);
    double* times = prog_ptr<double>(11);
    int* sc = (int*)gpumpi_malloc(sizeof(int) * 4 * (uint64_t)P);
    int* sd = sc + P;
    int* rc = sd + P;
    int* rd = rc + P;
    for (int i = 0; i < P; i++) {
        sc[i] = counts[(uint64_t)rank * P + i];
        rc[i] = counts[(uint64_t)i * P + rank];
        sd[i] = i == 0 ? 0 : sd[i - 1] + sc[i - 1];
        rd[i] = i == 0 ? 0 : rd[i - 1] + rc[i - 1];
    }
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    for (int it = 0; it < iters; it++) MPI_Alltoallv(sb, sc, sd, dt, rb, rc, rd, dt, MPI_COMM_WORLD);
    const double t1 = MPI_Wtime();
    if (times && gpumpi_is_leader_lane()) times[rank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

// Gather(v)/Scatter(v)/Allgather(v)/Alltoall(v) over int32 blocks.
// args: [1] kind (0 gather, 1 gatherv, 2 scatter, 3 scatterv, 4 allgather,
// 5 allgatherv, 6 alltoall, 7 alltoall in place, 8 alltoallv in place,
// 9 alltoall whose send blocks exceed the receive blocks), [2] block
// elements (int), [3] root, [4] send base, [5] recv base, [6] send stride,
// [7] recv stride.
// For the v-variants, rank r contributes (block + r) elements; the in-place
// Alltoallv exchanges block + (r + k) % 3 elements between ranks r and k
// (a symmetric count matrix, as MPI requires when sending in place).
GPUMPI_PROGRAM(coll_blocks) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int kind = (int)prog_int(1), blk = (int)prog_int(2), root = (int)prog_int(3);
    int* sb = (int*)(prog_ptr<uint8_t>(4) + (uint64_t)rank * prog_args()[6]);
    int* rb = (int*)(prog_ptr<uint8_t>(5) + (uint64_t)rank * prog_args()[7]);
    int* cnt = (int*)gpumpi_malloc(sizeof(int) * 2 * (uint64_t)P);
    int* dsp = cnt + P;
    for (int i = 0; i < P; i++) {
        cnt[i] = kind == 8 ? blk + (rank + i) % 3 : blk + i;
        dsp[i] = i == 0 ? 0 : dsp[i - 1] + cnt[i - 1];
    }
    switch (kind) {
    case 0: MPI_Gather(sb, blk, MPI_INT, rb, blk, MPI_INT, root, MPI_COMM_WORLD); break;
    case 1: MPI_Gatherv(sb, blk + rank, MPI_INT, rb, cnt, dsp, MPI_INT, root, MPI_COMM_WORLD); break;
    case 2: MPI_Scatter(sb, blk, MPI_INT, rb, blk, MPI_INT, root, MPI_COMM_WORLD); break;
    case 3: MPI_Scatterv(sb, cnt, dsp, MPI_INT, rb, blk + rank, MPI_INT, root, MPI_COMM_WORLD); break;
    case 4: MPI_Allgather(sb, blk, MPI_INT, rb, blk, MPI_INT, MPI_COMM_WORLD); break;
    case 5: MPI_Allgatherv(sb, blk + rank, MPI_INT, rb, cnt, dsp, MPI_INT, MPI_COMM_WORLD); break;
    case 6: MPI_Alltoall(sb, blk, MPI_INT, rb, blk, MPI_INT, MPI_COMM_WORLD); break;
    case 7: MPI_Alltoall(MPI_IN_PLACE, 0, MPI_INT, rb, blk, MPI_INT, MPI_COMM_WORLD); break;
    case 8: MPI_Alltoallv(MPI_IN_PLACE, nullptr, nullptr, MPI_INT, rb, cnt, dsp, MPI_INT, MPI_COMM_WORLD); break;
    default: MPI_Alltoall(sb, blk + 1, MPI_INT, rb, blk, MPI_INT, MPI_COMM_WORLD); break;
    }
    MPI_Finalize();
    return 0;
}

// The reference's toolchain_tests/pi/cpi.c as a rank program (converted by
// hand: main -> program, printf dropped).  args: [1] n, [2] out (double[2 +
// P]: pi at root, elapsed, then every rank's partial sum), [3] reduce only
// timing out (double[P]).
GPUMPI_PROGRAM(cpi) {
    int n, myid, numprocs, i;
    double mypi, pi = 0.0, h, sum, x;
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &numprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    double* out = prog_ptr<double>(2);
    double* rtimes = prog_ptr<double>(3);
    n = (int)prog_int(1);
    const double startwtime = MPI_Wtime();
    MPI_Bcast(&n, 1, MPI_INT, 0, MPI_COMM_WORLD);
    h = 1.0 / (double)n;
    sum = 0.0;
    for (i = myid + 1; i <= n; i += numprocs) {
        x = h * ((double)i - 0.5);
        sum += 4.0 / (1.0 + x * x);
    }
    mypi = h * sum;
    MPI_Barrier(MPI_COMM_WORLD);
    const double r0 = MPI_Wtime();
    MPI_Reduce(&mypi, &pi, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    const double r1 = MPI_Wtime();
    if (gpumpi_is_leader_lane()) {
        out[2 + myid] = mypi;
        if (rtimes) rtimes[myid] = r1 - r0;
    }
    if (myid == 0) {
        const double endwtime = MPI_Wtime();
        if (gpumpi_is_leader_lane()) {
            out[0] = pi;
            out[1] = endwtime - startwtime;
        }
    }
    MPI_Finalize();
    return 0;
}
