/*
 * Host-MPI golden vectors for the reduction operators (test infrastructure).
 *
 * Runs MPI_Reduce(root 0) for every (datatype, op) pair the reference's
 * legality table admits (gpu_libs/gpu_mpi/operators.cu:153-200) on the
 * datatypes of gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33, with the host MPI
 * library's own datatypes and ops, and writes rank 0's results to a binary
 * file.  The inputs are the integer-valued patterns below, so every result is
 * exact whatever order the host library combines ranks in.
 *
 * Record layout (little endian): int32 dtype, int32 op, int32 count,
 * int32 nbytes, then nbytes of result.
 */
#include <complex.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define COUNT 14

/* Our enum index -> host MPI datatype. The reference maps long double to an
 * 8-byte double (mpi_types_list.cuh:18,31), so the host run uses MPI_DOUBLE
 * and MPI_C_DOUBLE_COMPLEX for those two entries. */
static MPI_Datatype host_type(int t) {
    switch (t) {
    case 0: return MPI_CHAR; case 1: return MPI_SHORT; case 2: return MPI_INT;
    case 3: return MPI_LONG; case 4: return MPI_LONG_LONG_INT; case 5: return MPI_LONG_LONG;
    case 6: return MPI_SIGNED_CHAR; case 7: return MPI_UNSIGNED_CHAR;
    case 8: return MPI_UNSIGNED_SHORT; case 9: return MPI_UNSIGNED;
    case 10: return MPI_UNSIGNED_LONG; case 11: return MPI_UNSIGNED_LONG_LONG;
    case 12: return MPI_FLOAT; case 13: return MPI_DOUBLE; case 14: return MPI_DOUBLE;
    case 15: return MPI_WCHAR; case 16: return MPI_C_BOOL; case 17: return MPI_INT8_T;
    case 18: return MPI_INT16_T; case 19: return MPI_INT32_T; case 20: return MPI_INT64_T;
    case 21: return MPI_UINT8_T; case 22: return MPI_UINT16_T; case 23: return MPI_UINT32_T;
    case 24: return MPI_UINT64_T; case 25: return MPI_C_COMPLEX;
    case 26: return MPI_C_FLOAT_COMPLEX; case 27: return MPI_C_DOUBLE_COMPLEX;
    case 28: return MPI_C_DOUBLE_COMPLEX; case 29: return MPI_BYTE;
    }
    return MPI_DATATYPE_NULL;
}

static MPI_Op host_op(int op) {
    switch (op) {
    case 0: return MPI_MAX; case 1: return MPI_MIN; case 2: return MPI_SUM;
    case 3: return MPI_PROD; case 4: return MPI_LAND; case 5: return MPI_BAND;
    case 6: return MPI_LOR; case 7: return MPI_BOR; case 8: return MPI_LXOR;
    case 9: return MPI_BXOR;
    }
    return MPI_OP_NULL;
}

static const int k_size[30] = {1, 2, 4, 8, 8, 8, 1, 1, 2, 4, 8, 8, 4, 8, 8, 4, 1,
                               1, 2, 4, 8, 1, 2, 4, 8, 8, 8, 16, 16, 1};

static int is_int(int t) { return (t >= 1 && t <= 11) || (t >= 17 && t <= 24); }
static int is_unsigned(int t) {
    return t == 7 || t == 8 || t == 9 || t == 10 || t == 11 || (t >= 21 && t <= 24);
}
static int is_fp(int t) { return t >= 12 && t <= 14; }
static int is_cx(int t) { return t >= 25 && t <= 28; }

static int allowed(int t, int op) {
    switch (op) {
    case 0: case 1: return is_int(t) || is_fp(t);
    case 2: case 3: return is_int(t) || is_fp(t) || is_cx(t);
    case 4: case 6: case 8: return is_int(t) || t == 16;
    case 5: case 7: case 9: return is_int(t) || t == 29;
    }
    return 0;
}

/* Input patterns (the pytest side rebuilds the same values). */
static long long sval(int r, int i) { return (long long)((r * 7 + i * 5) % 13) - 6; }
static long long uval(int r, int i) { return (long long)((r * 7 + i * 5) % 13); }
static int bval(int r, int i) { return ((r + i) % 3) == 0; }
static int byteval(int r, int i) { return (r * 37 + i * 11) & 0xff; }
static long long imval(int r, int i) { return (long long)((r * 3 + i * 2) % 7) - 3; }

static void fill(int t, int r, void* buf) {
    for (int i = 0; i < COUNT; i++) {
        long long v = is_unsigned(t) ? uval(r, i) : sval(r, i);
        switch (k_size[t]) {
        default: break;
        }
        if (t == 16) { ((unsigned char*)buf)[i] = (unsigned char)bval(r, i); continue; }
        if (t == 29 || t == 0) { ((unsigned char*)buf)[i] = (unsigned char)byteval(r, i); continue; }
        if (t == 12) { ((float*)buf)[i] = (float)sval(r, i); continue; }
        if (t == 13 || t == 14) { ((double*)buf)[i] = (double)sval(r, i); continue; }
        if (t == 25 || t == 26) {
            ((float*)buf)[2 * i] = (float)sval(r, i);
            ((float*)buf)[2 * i + 1] = (float)imval(r, i);
            continue;
        }
        if (t == 27 || t == 28) {
            ((double*)buf)[2 * i] = (double)sval(r, i);
            ((double*)buf)[2 * i + 1] = (double)imval(r, i);
            continue;
        }
        switch (k_size[t]) {
        case 1: ((int8_t*)buf)[i] = (int8_t)v; break;
        case 2: ((int16_t*)buf)[i] = (int16_t)v; break;
        case 4: ((int32_t*)buf)[i] = (int32_t)v; break;
        case 8: ((int64_t*)buf)[i] = (int64_t)v; break;
        }
    }
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    FILE* f = NULL;
    if (rank == 0) {
        f = fopen(argc > 1 ? argv[1] : "reduce_golden.bin", "wb");
        if (!f) MPI_Abort(MPI_COMM_WORLD, 1);
    }
    unsigned char in[COUNT * 16], out[COUNT * 16];
    for (int t = 0; t < 30; t++) {
        for (int op = 0; op < 10; op++) {
            if (!allowed(t, op)) continue;
            memset(in, 0, sizeof in);
            memset(out, 0, sizeof out);
            fill(t, rank, in);
            MPI_Reduce(in, out, COUNT, host_type(t), host_op(op), 0, MPI_COMM_WORLD);
            if (rank == 0) {
                int hdr[4] = {t, op, COUNT, COUNT * k_size[t]};
                fwrite(hdr, sizeof hdr, 1, f);
                fwrite(out, 1, (size_t)hdr[3], f);
            }
        }
    }
    if (rank == 0) fclose(f);
    MPI_Finalize();
    return 0;
}
