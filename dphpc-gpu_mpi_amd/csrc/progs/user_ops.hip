// Reductions with user-defined operators (MPI_Op_create), checked in the
// program against closed forms and against the predefined operator:
//   * a user FP64 sum, inout[i] = in[i] + inout[i], equals MPI_SUM bit for
//     bit (the same rank-order fold, mpi.cu:209-225);
//   * an asymmetric int64 operator, inout[i] = 2 * in[i] + inout[i], gives
//     x_0 + 2 * (x_1 + ... + x_{P-1}): the fold starts from rank 0's input
//     and applies fn(x_r, acc) for r = 1 .. P-1 (the reference's order);
//   * MPI_Op_free returns the operator's storage.
// args: [0] ok (int[P]), [1] count, [2] root (-1: MPI_Allreduce),
//       [3] in place (MPI_IN_PLACE at the receiving ranks), [4] int64 out:
//       rank 0's (or the root's) asymmetric result, count values.
#include "registry.hpp"

namespace {

__device__ void user_sum_f64(void* in, void* inout, int* len, MPI_Datatype* dt) {
    const double* a = (const double*)in;
    double* b = (double*)inout;
    if (*dt != MPI_DOUBLE) return;
    for (int i = 0; i < *len; i++) b[i] = a[i] + b[i];
}

__device__ void user_twice_plus_i64(void* in, void* inout, int* len, MPI_Datatype* dt) {
    const long long* a = (const long long*)in;
    long long* b = (long long*)inout;
    (void)dt;
    for (int i = 0; i < *len; i++) b[i] = 2 * a[i] + b[i];
}

__device__ inline double value_f64(int rank, int i) {
    // spread over many binades so that the fold's order shows in the bits
    const uint64_t z = (uint64_t)(rank + 1) * 0x9E3779B97F4A7C15ULL + (uint64_t)i * 0xBF58476D1CE4E5B9ULL;
    const double m = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    return m * (double)(1ull << ((z >> 3) % 40));
}
__device__ inline long long value_i64(int rank, int i) {
    return ((long long)rank * 7919 + (long long)i * 104729) % 100003 - 50000;
}

}  // namespace

GPUMPI_PROGRAM(user_op_reduce) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int n = (int)prog_int(1), root = (int)prog_int(2);
    const bool in_place = prog_int(3) != 0;
    long long* out = prog_ptr<long long>(4);
    const bool receives = root < 0 || rank == root;
    const int lane = gpumpi::rt::lane_id();

    double* xd = (double*)gpumpi_malloc(sizeof(double) * (uint64_t)n);
    double* yd = (double*)gpumpi_malloc(sizeof(double) * (uint64_t)n);
    double* zd = (double*)gpumpi_malloc(sizeof(double) * (uint64_t)n);
    long long* xi = (long long*)gpumpi_malloc(sizeof(long long) * (uint64_t)n);
    long long* yi = (long long*)gpumpi_malloc(sizeof(long long) * (uint64_t)n);
    double* od = (double*)gpumpi_malloc(sizeof(double) * (uint64_t)n);
    if (!xd || !yd || !zd || !xi || !yi || !od) {
        MPI_Abort(MPI_COMM_WORLD, 98);
        return 1;
    }
    for (int i = lane; i < n; i += 64) {
        xd[i] = value_f64(rank, i);
        yd[i] = xd[i];
        zd[i] = 0.0;
        xi[i] = value_i64(rank, i);
        yi[i] = xi[i];
    }
    gpumpi::rt::wave_sync();
    // yd keeps this rank's input for the predefined MPI_SUM beside the user
    // sum; od and yi take results out of place

    MPI_Op usum, utwice;
    MPI_Op_create(user_sum_f64, 1, &usum);
    MPI_Op_create(user_twice_plus_i64, 0, &utwice);
    const void* sd = in_place && receives ? MPI_IN_PLACE : (const void*)xd;
    const void* si = in_place && receives ? MPI_IN_PLACE : (const void*)xi;
    double* rd = in_place && receives ? xd : od;
    long long* ri = in_place && receives ? xi : yi;
    if (root < 0) {
        MPI_Allreduce(sd, rd, n, MPI_DOUBLE, usum, MPI_COMM_WORLD);
        MPI_Allreduce(yd, zd, n, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
        MPI_Allreduce(si, ri, n, MPI_LONG_LONG, utwice, MPI_COMM_WORLD);
    } else {
        MPI_Reduce(sd, rd, n, MPI_DOUBLE, usum, root, MPI_COMM_WORLD);
        MPI_Reduce(yd, zd, n, MPI_DOUBLE, MPI_SUM, root, MPI_COMM_WORLD);
        MPI_Reduce(si, ri, n, MPI_LONG_LONG, utwice, root, MPI_COMM_WORLD);
    }
    bool ok = true;
    if (receives) {
        bool bad = false;
        for (int i = lane; i < n; i += 64) {
            long long want = 0;
            for (int r = 1; r < P; r++) want += value_i64(r, i);
            want = value_i64(0, i) + 2 * want;
            if (ri[i] != want) bad = true;
            if (__builtin_bit_cast(uint64_t, rd[i]) != __builtin_bit_cast(uint64_t, zd[i])) bad = true;
        }
        ok = gpumpi::rt::ballot(bad) == 0;
        if (out && rank == (root < 0 ? 0 : root))
            for (int i = lane; i < n; i += 64) out[i] = ri[i];
        gpumpi::rt::wave_sync();
    }
    MPI_Op_free(&usum);
    MPI_Op_free(&utwice);
    ok &= usum == MPI_OP_NULL && utwice == MPI_OP_NULL;
    MPI_Finalize();
    if (gpumpi_is_leader_lane()) prog_ptr<int>(0)[rank] = ok ? 1 : 0;
    return 0;
}
