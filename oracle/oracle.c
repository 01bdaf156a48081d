/*
 * CPU oracle for the device-resident MPI hot path -- TEST INFRASTRUCTURE.
 * See oracle.h for what each function restates and how it is pinned.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off; no FMA contraction, so
 * FP64 results are the plain IEEE sequence the reference's C code describes).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------
 * Datatypes: gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33 (enum order, sizes).
 * ---------------------------------------------------------------------- */
enum {
    T_CHAR, T_SHORT, T_INT, T_LONG, T_LONG_LONG_INT, T_LONG_LONG,
    T_SIGNED_CHAR, T_UNSIGNED_CHAR, T_UNSIGNED_SHORT, T_UNSIGNED,
    T_UNSIGNED_LONG, T_UNSIGNED_LONG_LONG, T_FLOAT, T_DOUBLE, T_LONG_DOUBLE,
    T_WCHAR, T_C_BOOL, T_INT8, T_INT16, T_INT32, T_INT64, T_UINT8, T_UINT16,
    T_UINT32, T_UINT64, T_C_COMPLEX, T_C_FLOAT_COMPLEX, T_C_DOUBLE_COMPLEX,
    T_C_LONG_DOUBLE_COMPLEX, T_BYTE, T_COUNT
};

/* Operators: gpu_libs/gpu_mpi/mpi_operators_list.cuh:4-16 (enum order). */
enum { O_MAX, O_MIN, O_SUM, O_PROD, O_LAND, O_BAND, O_LOR, O_BOR, O_LXOR,
       O_BXOR, O_MAXLOC, O_MINLOC, O_COUNT };

static const int k_type_size[T_COUNT] = {
    1, 2, 4, 8, 8, 8, 1, 1, 2, 4, 8, 8, 4, 8, 8, 4, 1,
    1, 2, 4, 8, 1, 2, 4, 8, 8, 8, 16, 16, 1};

int orc_type_size(int dtype) {
    if (dtype < 0 || dtype >= T_COUNT) return -1;
    return k_type_size[dtype];
}

/* Type classes of operators.cu:25-71. */
static int is_integer(int t) {
    switch (t) {
    case T_SHORT: case T_INT: case T_LONG: case T_LONG_LONG_INT:
    case T_LONG_LONG: case T_SIGNED_CHAR: case T_UNSIGNED_CHAR:
    case T_UNSIGNED_SHORT: case T_UNSIGNED: case T_UNSIGNED_LONG:
    case T_UNSIGNED_LONG_LONG: case T_INT8: case T_INT16: case T_INT32:
    case T_INT64: case T_UINT8: case T_UINT16: case T_UINT32: case T_UINT64:
        return 1;
    default: return 0;
    }
}
static int is_floating(int t) { return t == T_FLOAT || t == T_DOUBLE || t == T_LONG_DOUBLE; }
static int is_logical(int t) { return t == T_C_BOOL; }
static int is_complex(int t) {
    return t == T_C_COMPLEX || t == T_C_FLOAT_COMPLEX || t == T_C_DOUBLE_COMPLEX ||
           t == T_C_LONG_DOUBLE_COMPLEX;
}
static int is_byte(int t) { return t == T_BYTE; }

/* Legality table, operators.cu:153-200. MAXLOC/MINLOC have no entry there,
 * so no built-in datatype admits them. */
int orc_op_allowed(int t, int op) {
    if (t < 0 || t >= T_COUNT) return 0;
    switch (op) {
    case O_MAX: case O_MIN: return is_integer(t) || is_floating(t);
    case O_SUM: case O_PROD: return is_integer(t) || is_floating(t) || is_complex(t);
    case O_LAND: case O_LOR: case O_LXOR: return is_integer(t) || is_logical(t);
    case O_BAND: case O_BOR: case O_BXOR: return is_integer(t) || is_byte(t);
    default: return 0;
    }
}

/* One step acc = op(in, acc) of operators.cu:216-218 for an integer type T
 * (unsigned U used for wrap-around arithmetic, as two's-complement hardware
 * does).  LXOR is the MPI logical exclusive or. */
#define INT_FOLD(T, U)                                                         \
    static void fold_##T(int op, size_t n, const T* in, T* acc) {              \
        for (size_t i = 0; i < n; i++) {                                       \
            T a = in[i], b = acc[i], c;                                        \
            switch (op) {                                                      \
            case O_MAX: c = a > b ? a : b; break;                              \
            case O_MIN: c = a > b ? b : a; break;                              \
            case O_SUM: c = (T)((U)a + (U)b); break;                           \
            case O_PROD: c = (T)((U)a * (U)b); break;                          \
            case O_LAND: c = (T)(a && b); break;                               \
            case O_LOR: c = (T)(a || b); break;                                \
            case O_LXOR: c = (T)((a != 0) != (b != 0)); break;                 \
            case O_BAND: c = (T)(a & b); break;                                \
            case O_BOR: c = (T)(a | b); break;                                 \
            case O_BXOR: c = (T)(a ^ b); break;                                \
            default: c = b; break;                                             \
            }                                                                  \
            acc[i] = c;                                                        \
        }                                                                      \
    }

typedef signed char schar;
typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef unsigned long ulong;
typedef long long llong;
typedef unsigned long long ullong;
INT_FOLD(schar, uchar)
INT_FOLD(uchar, uchar)
INT_FOLD(short, ushort)
INT_FOLD(ushort, ushort)
INT_FOLD(int, uint)
INT_FOLD(uint, uint)
INT_FOLD(llong, ullong)
INT_FOLD(ullong, ullong)

#define FP_FOLD(T)                                                             \
    static void fold_##T(int op, size_t n, const T* in, T* acc) {              \
        for (size_t i = 0; i < n; i++) {                                       \
            T a = in[i], b = acc[i], c;                                        \
            switch (op) {                                                      \
            case O_MAX: c = a > b ? a : b; break;                              \
            case O_MIN: c = a > b ? b : a; break;                              \
            case O_SUM: c = a + b; break;                                      \
            case O_PROD: c = a * b; break;                                     \
            default: c = b; break;                                             \
            }                                                                  \
            acc[i] = c;                                                        \
        }                                                                      \
    }
FP_FOLD(float)
FP_FOLD(double)

/* Complex SUM/PROD on (re, im) pairs: (a*b) = (ar*br - ai*bi, ar*bi + ai*br),
 * the textbook product without C99 Annex G NaN recovery. */
#define CX_FOLD(T)                                                             \
    static void foldcx_##T(int op, size_t n, const T* in, T* acc) {            \
        for (size_t i = 0; i < n; i++) {                                       \
            T ar = in[2 * i], ai = in[2 * i + 1];                              \
            T br = acc[2 * i], bi = acc[2 * i + 1];                            \
            if (op == O_SUM) {                                                 \
                acc[2 * i] = ar + br;                                          \
                acc[2 * i + 1] = ai + bi;                                      \
            } else {                                                           \
                T re = ar * br - ai * bi;                                      \
                T im = ar * bi + ai * br;                                      \
                acc[2 * i] = re;                                               \
                acc[2 * i + 1] = im;                                           \
            }                                                                  \
        }                                                                      \
    }
CX_FOLD(float)
CX_FOLD(double)

static void fold_step(int t, int op, size_t n, const void* in, void* acc) {
    switch (t) {
    case T_SIGNED_CHAR: case T_INT8: fold_schar(op, n, in, acc); break;
    case T_UNSIGNED_CHAR: case T_UINT8: case T_C_BOOL: case T_BYTE:
        fold_uchar(op, n, in, acc); break;
    case T_SHORT: case T_INT16: fold_short(op, n, in, acc); break;
    case T_UNSIGNED_SHORT: case T_UINT16: fold_ushort(op, n, in, acc); break;
    case T_INT: case T_INT32: fold_int(op, n, in, acc); break;
    case T_UNSIGNED: case T_UINT32: fold_uint(op, n, in, acc); break;
    case T_LONG: case T_LONG_LONG_INT: case T_LONG_LONG: case T_INT64:
        fold_llong(op, n, in, acc); break;
    case T_UNSIGNED_LONG: case T_UNSIGNED_LONG_LONG: case T_UINT64:
        fold_ullong(op, n, in, acc); break;
    case T_FLOAT: fold_float(op, n, in, acc); break;
    case T_DOUBLE: case T_LONG_DOUBLE: fold_double(op, n, in, acc); break;
    case T_C_COMPLEX: case T_C_FLOAT_COMPLEX: foldcx_float(op, n, in, acc); break;
    case T_C_DOUBLE_COMPLEX: case T_C_LONG_DOUBLE_COMPLEX:
        foldcx_double(op, n, in, acc); break;
    default: break;
    }
}

int orc_reduce_fold(int dtype, int op, int nranks, size_t count,
                    const void* inputs, void* out) {
    if (!orc_op_allowed(dtype, op) || nranks <= 0) return -1;
    size_t bytes = (size_t)k_type_size[dtype] * count;
    const char* base = (const char*)inputs;
    memcpy(out, base, bytes); /* src 0 is copied (mpi.cu:218-221) */
    for (int r = 1; r < nranks; r++)
        fold_step(dtype, op, count, base + (size_t)r * bytes, out);
    return 0;
}

/* ------------------------------------------------------------------------
 * pi: toolchain_tests/pi/cpi.c:15-18 (f) and :47-53 (cyclic partial sum).
 * ---------------------------------------------------------------------- */
double orc_pi_partial(int n, int rank, int nranks) {
    double h = 1.0 / (double)n;
    double sum = 0.0;
    for (int i = rank + 1; i <= n; i += nranks) {
        double x = h * ((double)i - 0.5);
        sum += 4.0 / (1.0 + x * x);
    }
    return h * sum;
}

double orc_pi_fold(int n, int nranks, double* partials_out) {
    double acc = 0.0;
    for (int r = 0; r < nranks; r++) {
        double p = orc_pi_partial(n, r, nranks);
        if (partials_out) partials_out[r] = p;
        acc = (r == 0) ? p : p + acc; /* acc = op(x_r, acc), mpi.cu:218-224 */
    }
    return acc;
}

/* ------------------------------------------------------------------------
 * Seeded payloads.  mix64 is the splitmix64 output function (Steele, Lea,
 * Flood 2014); the word generator feeds it a Weyl sequence per (seed, rank).
 * ---------------------------------------------------------------------- */
uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_splitmix64_next(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ULL;
    return orc_mix64(*state);
}

uint64_t orc_payload_word(uint64_t seed, uint64_t rank, uint64_t i) {
    return orc_mix64(seed + rank * 0xD1B54A32D192ED03ULL + (i + 1) * 0x9E3779B97F4A7C15ULL);
}

void orc_payload_bytes(uint64_t seed, uint64_t rank, size_t nbytes, uint8_t* out) {
    size_t nw = nbytes / 8;
    for (size_t i = 0; i < nw; i++) {
        uint64_t w = orc_payload_word(seed, rank, i);
        memcpy(out + 8 * i, &w, 8);
    }
    if (nbytes % 8) {
        uint64_t w = orc_payload_word(seed, rank, nw);
        memcpy(out + 8 * nw, &w, nbytes % 8);
    }
}

/* ------------------------------------------------------------------------
 * NPB IS.  Algorithm of toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c,
 * restated for P simulated ranks in one process.
 * ---------------------------------------------------------------------- */
#define IS_MAX_ITERATIONS 10
#define IS_TEST_ARRAY_SIZE 5

/* randlc, is.c:408-476: x_{k+1} = a x_k mod 2^46, returned as 2^-46 x. */
double orc_randlc(double* x, double a) {
    const double r23 = 1.0 / 8388608.0, r46 = r23 * r23;
    const double t23 = 8388608.0, t46 = t23 * t23;
    double t1, t2, t3, t4, a1, a2, x1, x2, z;
    int j;
    t1 = r23 * a;
    j = (int)t1;
    a1 = j;
    a2 = a - t23 * a1;
    t1 = r23 * *x;
    j = (int)t1;
    x1 = j;
    x2 = *x - t23 * x1;
    t1 = a1 * x2 + a2 * x1;
    j = (int)(r23 * t1);
    t2 = j;
    z = t1 - t23 * t2;
    t3 = t23 * z + a2 * x2;
    j = (int)(r46 * t3);
    t4 = j;
    *x = t3 - t46 * t4;
    return r46 * *x;
}

/* find_my_seed, is.c:489-533: jump the LCG to rank kn's first number. */
double orc_find_my_seed(int kn, int np, long nn, double s, double a) {
    long nq = nn / np, mq, kk, ik, i;
    double t1, t2, an;
    for (mq = 0; nq > 1; mq++, nq /= 2)
        ;
    t1 = a;
    for (i = 1; i <= mq; i++) orc_randlc(&t1, t1);
    an = t1;
    kk = kn;
    t1 = s;
    t2 = an;
    for (i = 1; i <= 100; i++) {
        ik = kk / 2;
        if (2 * ik != kk) orc_randlc(&t1, t2);
        if (ik == 0) break;
        orc_randlc(&t2, t2);
        kk = ik;
    }
    return t1;
}

typedef struct is_params {
    int total_keys_log2, max_key_log2, num_buckets_log2;
    const int* test_index;
    const int* test_rank;
    char cls;
} is_params;

/* Known-answer arrays, is.c:254-274. */
static const int S_idx[5] = {48427, 17148, 23627, 62548, 4431};
static const int S_rnk[5] = {0, 18, 346, 64917, 65463};
static const int W_idx[5] = {357773, 934767, 875723, 898999, 404505};
static const int W_rnk[5] = {1249, 11698, 1039987, 1043896, 1048018};
static const int A_idx[5] = {2112377, 662041, 5336171, 3642833, 4250760};
static const int A_rnk[5] = {104, 17523, 123928, 8288932, 8388264};
static const int B_idx[5] = {41869, 812306, 5102857, 18232239, 26860214};
static const int B_rnk[5] = {33422937, 10244, 59149, 33135281, 99};

/* Class table, is.c:60-95. */
static int is_get_params(char cls, is_params* p) {
    p->cls = cls;
    switch (cls) {
    case 'S': p->total_keys_log2 = 16; p->max_key_log2 = 11; p->num_buckets_log2 = 9;
        p->test_index = S_idx; p->test_rank = S_rnk; return 0;
    case 'W': p->total_keys_log2 = 20; p->max_key_log2 = 16; p->num_buckets_log2 = 10;
        p->test_index = W_idx; p->test_rank = W_rnk; return 0;
    case 'A': p->total_keys_log2 = 23; p->max_key_log2 = 19; p->num_buckets_log2 = 10;
        p->test_index = A_idx; p->test_rank = A_rnk; return 0;
    case 'B': p->total_keys_log2 = 25; p->max_key_log2 = 21; p->num_buckets_log2 = 10;
        p->test_index = B_idx; p->test_rank = B_rnk; return 0;
    default: return -1;
    }
}

/* Per-iteration adjustment of the expected rank, is.c:869-913. */
static long is_expected_rank(char cls, int i, int iteration, long test_rank) {
    switch (cls) {
    case 'S': return (i <= 2) ? test_rank + iteration : test_rank - iteration;
    case 'W': return (i < 2) ? test_rank + iteration - 2 : test_rank - iteration;
    case 'A': return (i <= 2) ? test_rank + iteration - 1 : test_rank - (iteration - 1);
    case 'B': return (i == 1 || i == 2 || i == 4) ? test_rank + iteration
                                                  : test_rank - iteration;
    default: return test_rank;
    }
}

typedef struct is_rank_state {
    int32_t* key_array;
    int32_t* key_buff1; /* bucketed keys, then the per-key counts */
    int32_t* key_buff2; /* received keys */
    int* bucket_size;   /* NUM_BUCKETS + 5 */
    int* send_count;
    int* send_displ;
    int* recv_count;
    int* recv_displ;
    int total_local_keys;
    int min_key_val;
    int32_t* counts; /* key_buff_ptr + min_key_val of the last iteration */
    int counts_len;
} is_rank_state;

static uint64_t is_checksum(const int32_t* keys, size_t n, size_t offset) {
    uint64_t s = 0;
    for (size_t i = 0; i < n; i++)
        s += orc_mix64((uint64_t)(uint32_t)keys[i] + ((uint64_t)(offset + i) << 32));
    return s;
}

int orc_is_keys(char cls, int nranks, int rank, int32_t* out) {
    is_params p;
    if (is_get_params(cls, &p)) return -1;
    long total_keys = 1L << p.total_keys_log2;
    int num_keys = (int)(total_keys / nranks);
    double seed = orc_find_my_seed(rank, nranks, 4 * total_keys, 314159265.00, 1220703125.00);
    double a = 1220703125.00;
    int k = (1 << p.max_key_log2) / 4;
    for (int i = 0; i < num_keys; i++) { /* create_seq, is.c:535-553 */
        double x = orc_randlc(&seed, a);
        x += orc_randlc(&seed, a);
        x += orc_randlc(&seed, a);
        x += orc_randlc(&seed, a);
        out[i] = (int32_t)(k * x);
    }
    return 0;
}

int orc_is_run(char cls, int nranks, orc_is_result* res, int32_t* keys_out,
               int32_t* bucket_totals_out) {
    is_params p;
    if (is_get_params(cls, &p) || nranks <= 0 || (nranks & (nranks - 1))) return -1;
    const int P = nranks;
    const long total_keys = 1L << p.total_keys_log2;
    const int max_key = 1 << p.max_key_log2;
    const int nb = 1 << p.num_buckets_log2;
    const int shift = p.max_key_log2 - p.num_buckets_log2;
    const int num_keys = (int)(total_keys / P);
    int size_of_buffers; /* alloc_space, is.c:333-345 */
    if (P < 256) size_of_buffers = 3 * num_keys / 2;
    else if (P < 512) size_of_buffers = 5 * num_keys / 2;
    else if (P < 1024) size_of_buffers = 4 * num_keys;
    else size_of_buffers = 13 * num_keys / 2;

    is_rank_state* st = calloc(P, sizeof(is_rank_state));
    int* totals = calloc(nb + IS_TEST_ARRAY_SIZE, sizeof(int));
    int* ptr1 = calloc(nb + IS_TEST_ARRAY_SIZE, sizeof(int));
    int* ptr2 = calloc(nb + IS_TEST_ARRAY_SIZE, sizeof(int));
    int* bucket_ptrs = calloc(nb, sizeof(int));
    if (!st || !totals || !ptr1 || !ptr2 || !bucket_ptrs) return -2;
    for (int r = 0; r < P; r++) {
        st[r].key_array = calloc(size_of_buffers, 4);
        st[r].key_buff1 = calloc(size_of_buffers, 4);
        st[r].key_buff2 = calloc(size_of_buffers, 4);
        st[r].bucket_size = calloc(nb + IS_TEST_ARRAY_SIZE, sizeof(int));
        st[r].send_count = calloc(P, sizeof(int));
        st[r].send_displ = calloc(P, sizeof(int));
        st[r].recv_count = calloc(P, sizeof(int));
        st[r].recv_displ = calloc(P, sizeof(int));
        if (!st[r].key_array || !st[r].key_buff1 || !st[r].key_buff2) return -2;
        orc_is_keys(cls, P, r, st[r].key_array);
    }

    int passed = 0;
    struct timespec t_begin = {0, 0}, t_end = {0, 0};
    for (int it_index = 0; it_index <= IS_MAX_ITERATIONS; it_index++) {
        /* it_index 0 is the untimed rank(1) of is.c:1075; the counter is reset
         * after it (is.c:1078); the timer brackets the 10 timed iterations as
         * is.c:1090-1101 does. */
        int iteration = it_index == 0 ? 1 : it_index;
        if (it_index == 1) {
            passed = 0;
            clock_gettime(CLOCK_MONOTONIC, &t_begin);
        }

        /* rank(): local part before the Allreduce, is.c:648-690. */
        for (int r = 0; r < P; r++) {
            is_rank_state* s = &st[r];
            if (r == 0) {
                s->key_array[iteration] = iteration;
                s->key_array[iteration + IS_MAX_ITERATIONS] = max_key - iteration;
            }
            memset(s->bucket_size, 0, sizeof(int) * (nb + IS_TEST_ARRAY_SIZE));
            for (int i = 0; i < IS_TEST_ARRAY_SIZE; i++)
                if (p.test_index[i] / num_keys == r)
                    s->bucket_size[nb + i] = s->key_array[p.test_index[i] % num_keys];
            for (int i = 0; i < num_keys; i++) s->bucket_size[s->key_array[i] >> shift]++;
            bucket_ptrs[0] = 0;
            for (int i = 1; i < nb; i++) bucket_ptrs[i] = bucket_ptrs[i - 1] + s->bucket_size[i - 1];
            for (int i = 0; i < num_keys; i++) {
                int key = s->key_array[i];
                s->key_buff1[bucket_ptrs[key >> shift]++] = key;
            }
        }
        /* Allreduce(SUM, MPI_INT) of bucket_size, is.c:696-701. */
        memset(totals, 0, sizeof(int) * (nb + IS_TEST_ARRAY_SIZE));
        for (int r = 0; r < P; r++)
            for (int i = 0; i < nb + IS_TEST_ARRAY_SIZE; i++) totals[i] += st[r].bucket_size[i];

        /* Redistribution plan, is.c:706-751 (ptr1/ptr2 are identical on all
         * ranks because they depend on the totals only). */
        for (int r = 0; r < P; r++) {
            is_rank_state* s = &st[r];
            long bucket_sum_accumulator = 0;
            int local_acc = 0, j = 0;
            memset(ptr1, 0, sizeof(int) * (nb + IS_TEST_ARRAY_SIZE));
            memset(ptr2, 0, sizeof(int) * (nb + IS_TEST_ARRAY_SIZE));
            s->send_displ[0] = 0;
            ptr1[0] = 0;
            for (int i = 0; i < nb; i++) {
                bucket_sum_accumulator += totals[i];
                local_acc += s->bucket_size[i];
                if (bucket_sum_accumulator >= (long)(j + 1) * num_keys) {
                    s->send_count[j] = local_acc;
                    if (j != 0) {
                        s->send_displ[j] = s->send_displ[j - 1] + s->send_count[j - 1];
                        ptr1[j] = ptr2[j - 1] + 1;
                    }
                    ptr2[j++] = i;
                    local_acc = 0;
                }
            }
            while (j < P) {
                s->send_count[j] = 0;
                ptr1[j] = 1;
                j++;
            }
        }
        /* Alltoall of counts and Alltoallv of keys, is.c:758-781. */
        for (int r = 0; r < P; r++) {
            is_rank_state* s = &st[r];
            for (int src = 0; src < P; src++) s->recv_count[src] = st[src].send_count[r];
            s->recv_displ[0] = 0;
            for (int i = 1; i < P; i++) s->recv_displ[i] = s->recv_displ[i - 1] + s->recv_count[i - 1];
        }
        for (int r = 0; r < P; r++) {
            is_rank_state* s = &st[r];
            for (int src = 0; src < P; src++)
                memcpy(s->key_buff2 + s->recv_displ[src],
                       st[src].key_buff1 + st[src].send_displ[r],
                       sizeof(int32_t) * (size_t)s->recv_count[src]);
        }
        /* Local ranking and partial verification, is.c:787-913. */
        for (int r = 0; r < P; r++) {
            is_rank_state* s = &st[r];
            int min_key_val = ptr1[r] << shift;
            int max_key_val = ((ptr2[r] + 1) << shift) - 1;
            int range = max_key_val - min_key_val + 1;
            if (range < 1) range = 1;
            if (s->counts_len < range) {
                free(s->counts);
                s->counts = calloc(range, sizeof(int32_t));
                s->counts_len = range;
            }
            memset(s->counts, 0, sizeof(int32_t) * (size_t)range);
            long m = 0;
            for (int k = 0; k < r; k++)
                for (int i = ptr1[k]; i <= ptr2[k]; i++) m += totals[i];
            int j = 0;
            for (int i = ptr1[r]; i <= ptr2[r]; i++) j += totals[i];
            for (int i = 0; i < j; i++) s->counts[s->key_buff2[i] - min_key_val]++;
            for (int i = min_key_val; i < max_key_val; i++)
                s->counts[i + 1 - min_key_val] += s->counts[i - min_key_val];
            for (int i = 0; i < IS_TEST_ARRAY_SIZE; i++) {
                int k = totals[i + nb];
                if (min_key_val <= k && k <= max_key_val) {
                    long below = (k - 1 >= min_key_val) ? s->counts[k - 1 - min_key_val] : 0;
                    long key_rank = below + m;
                    long test_rank = is_expected_rank(p.cls, i, iteration, p.test_rank[i]);
                    if (key_rank == test_rank) passed++;
                }
            }
            if (it_index == IS_MAX_ITERATIONS) {
                s->total_local_keys = j;
                s->min_key_val = min_key_val;
            }
        }
        if (bucket_totals_out && it_index == IS_MAX_ITERATIONS)
            memcpy(bucket_totals_out, totals, sizeof(int) * (nb + IS_TEST_ARRAY_SIZE));
    }
    clock_gettime(CLOCK_MONOTONIC, &t_end);
    res->timed_seconds = (double)(t_end.tv_sec - t_begin.tv_sec) + 1e-9 * (double)(t_end.tv_nsec - t_begin.tv_nsec);

    /* full_verify, is.c:561-625. */
    int errors = 0;
    int64_t total_after = 0;
    uint64_t checksum = 0;
    size_t offset = 0;
    for (int r = 0; r < P; r++) {
        is_rank_state* s = &st[r];
        for (int i = 0; i < s->total_local_keys; i++) {
            int key = s->key_buff2[i];
            s->key_array[--s->counts[key - s->min_key_val]] = key;
        }
    }
    for (int r = 0; r < P; r++) {
        is_rank_state* s = &st[r];
        int j = 0;
        if (r > 0 && s->total_local_keys > 0) {
            const is_rank_state* prev = &st[r - 1];
            int last = prev->total_local_keys < 1 ? 0 : prev->total_local_keys - 1;
            if (prev->key_array[last] > s->key_array[0]) j++;
        }
        for (int i = 1; i < s->total_local_keys; i++)
            if (s->key_array[i - 1] > s->key_array[i]) j++;
        if (j == 0) passed++;
        errors += j;
        total_after += s->total_local_keys;
        checksum += is_checksum(s->key_array, (size_t)s->total_local_keys, offset);
        if (keys_out)
            memcpy(keys_out + offset, s->key_array, sizeof(int32_t) * (size_t)s->total_local_keys);
        offset += (size_t)s->total_local_keys;
    }
    res->passed_verification = passed;
    res->full_verify_errors = errors;
    res->total_keys_after = total_after;
    res->sorted_checksum = checksum;

    for (int r = 0; r < P; r++) {
        free(st[r].key_array); free(st[r].key_buff1); free(st[r].key_buff2);
        free(st[r].bucket_size); free(st[r].send_count); free(st[r].send_displ);
        free(st[r].recv_count); free(st[r].recv_displ); free(st[r].counts);
    }
    free(st); free(totals); free(ptr1); free(ptr2); free(bucket_ptrs);
    return 0;
}

/* ------------------------------------------------------------------------
 * BF16 and OCP FP8 e4m3 codecs, from the published formats.
 * ---------------------------------------------------------------------- */
uint16_t orc_f32_to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float orc_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

uint8_t orc_f32_to_e4m3(float xf) {
    uint32_t u;
    memcpy(&u, &xf, 4);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    double a = fabs((double)xf);
    if (xf != xf) return (uint8_t)(sign | 0x7f);
    if (a >= 448.0) return (uint8_t)(sign | 0x7e); /* saturate (no inf in e4m3fn) */
    if (a < 0.015625) {                             /* below 2^-6: subnormal grid 2^-9 */
        double q = rint(a * 512.0);                 /* round half to even */
        return (uint8_t)(sign | (uint8_t)q);        /* q == 8 is exactly 2^-6 */
    }
    int e;
    double m = frexp(a, &e); /* a = m * 2^e, m in [0.5, 1) */
    e -= 1;                  /* a = (2m) * 2^e with 2m in [1, 2) */
    double frac = rint((2.0 * m - 1.0) * 8.0);
    int mant = (int)frac;
    if (mant == 8) { mant = 0; e += 1; }
    int code = ((e + 7) << 3) | mant;
    if (code > 0x7e) code = 0x7e;
    return (uint8_t)(sign | code);
}

float orc_e4m3_to_f32(uint8_t q) {
    int s = q >> 7, e = (q >> 3) & 15, m = q & 7;
    float v;
    if (e == 15 && m == 7) v = NAN;
    else if (e == 0) v = (float)ldexp((double)m, -9);
    else v = (float)ldexp(1.0 + m / 8.0, e - 7);
    return s ? -v : v;
}

void orc_e4m3_block_encode(const float* x, size_t n, int block, uint8_t* q, float* scales) {
    size_t nblk = (n + (size_t)block - 1) / (size_t)block;
    for (size_t b = 0; b < nblk; b++) {
        size_t lo = b * (size_t)block, hi = lo + (size_t)block;
        if (hi > n) hi = n;
        float amax = 0.0f;
        for (size_t i = lo; i < hi; i++) {
            float v = fabsf(x[i]);
            if (v > amax) amax = v;
        }
        float scale = amax > 0.0f ? amax / 448.0f : 1.0f;
        float inv = amax > 0.0f ? 448.0f / amax : 1.0f;
        scales[b] = scale;
        for (size_t i = lo; i < hi; i++) q[i] = orc_f32_to_e4m3(x[i] * inv);
    }
}

void orc_e4m3_block_decode(const uint8_t* q, const float* scales, size_t n, int block, float* y) {
    for (size_t i = 0; i < n; i++) y[i] = orc_e4m3_to_f32(q[i]) * scales[i / (size_t)block];
}

/* ------------------------------------------------------------------------
 * Comm_split ordering, communicator.cu:147-213: members of one colour
 * ordered by key, ties by parent rank (rank_key_compare, :154-162).
 * ---------------------------------------------------------------------- */
void orc_comm_split(int nranks, const int* colors, const int* keys,
                    int* new_rank, int* new_size) {
    for (int r = 0; r < nranks; r++) {
        int rank = 0, size = 0;
        for (int q = 0; q < nranks; q++) {
            if (colors[q] != colors[r]) continue;
            size++;
            if (keys[q] < keys[r] || (keys[q] == keys[r] && q < r)) rank++;
        }
        new_rank[r] = rank;
        new_size[r] = size;
    }
}

/* ------------------------------------------------------------------------
 * 2-D particle-in-cell cycle (the C5 workload), decomposed over a px x py
 * grid of ranks exactly as the rank program csrc/progs/pic.hip runs it, and
 * the same physics on the undecomposed global grid.  The workload is this
 * build's own (the reference's sputniPIC, toolchain_tests/sputnipic/src/
 * mpi_comm.c:21-307, decomposes particles, not space); the oracle.h block
 * states every rule both sides follow.
 * ---------------------------------------------------------------------- */
static float pic_unit(uint64_t w) { return (float)(w >> 40) * (1.0f / 16777216.0f); }
static uint64_t pic_draw(uint64_t seed, int rank, int64_t i, int k) {
    return orc_mix64(seed + (uint64_t)rank * 0xD1B54A32D192ED03ULL +
                     (uint64_t)(i * 32 + k + 1) * 0x9E3779B97F4A7C15ULL);
}
void orc_pic_init_particle(const orc_pic_params* p, int rank, int64_t i, float* xyuv) {
    const int rx = rank % p->px, ry = rank / p->px;
    const float x0 = (float)(rx * p->nx), y0 = (float)(ry * p->ny);
    float x = x0 + (float)p->nx * pic_unit(pic_draw(p->seed, rank, i, 0));
    float y = y0 + (float)p->ny * pic_unit(pic_draw(p->seed, rank, i, 1));
    if (x >= x0 + (float)p->nx) x = x0;
    if (y >= y0 + (float)p->ny) y = y0;
    float gu = 0.0f, gv = 0.0f;
    for (int k = 0; k < 12; k++) gu = gu + pic_unit(pic_draw(p->seed, rank, i, 2 + k));
    for (int k = 0; k < 12; k++) gv = gv + pic_unit(pic_draw(p->seed, rank, i, 14 + k));
    xyuv[0] = x;
    xyuv[1] = y;
    xyuv[2] = p->drift + p->vth * (gu - 6.0f);
    xyuv[3] = p->vth * (gv - 6.0f);
}

/* direction index of (dx, dy) in {-1,0,1}^2 \ (0,0): 1..8, row by row */
static const int PIC_DX[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
static const int PIC_DY[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};

static void pic_push(const orc_pic_params* p, float* q, const float* ex, const float* ey, int w, int x0, int y0) {
    const float Lx = (float)(p->px * p->nx), Ly = (float)(p->py * p->ny);
    const int lx = (int)q[0] - x0, ly = (int)q[1] - y0;
    const float e1 = ex[(size_t)ly * w + lx], e2 = ey[(size_t)ly * w + lx];
    float u = q[2] + p->qm * e1;
    float v = q[3] + p->qm * e2;
    float x = q[0] + u, y = q[1] + v;
    if (x < 0.0f) x = x + Lx;
    if (x >= Lx) x = x - Lx;
    if (y < 0.0f) y = y + Ly;
    if (y >= Ly) y = y - Ly;
    q[0] = x;
    q[1] = y;
    q[2] = u;
    q[3] = v;
}

/* phi over the interior plus one ring, then E = -grad(phi) on the interior,
 * from rho with a two-cell ring; arrays (h+4) x (w+4), interior at [2, 2). */
static void pic_field(const float* rho, float* phi, float* ex, float* ey, int w, int h) {
    const int W = w + 4;
    for (int j = 1; j < h + 3; j++)
        for (int i = 1; i < w + 3; i++) {
            const float* r = rho + (size_t)j * W + i;
            phi[(size_t)j * W + i] = ((((4.0f * r[0] + r[-1]) + r[1]) + r[-W]) + r[W]) * 0.125f;
        }
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) {
            const float* f = phi + (size_t)(j + 2) * W + (i + 2);
            ex[(size_t)j * w + i] = (f[-1] - f[1]) * 0.5f;
            ey[(size_t)j * w + i] = (f[-W] - f[W]) * 0.5f;
        }
}

static uint64_t pic_hash(const void* data, size_t n) {
    const uint8_t* b = (const uint8_t*)data;
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ULL;
    return h;
}

/* energy of n particles: 1024-particle blocks summed in order */
double orc_pic_energy(const float* q, int64_t n) {
    double total = 0.0;
    for (int64_t b = 0; b < n; b += 1024) {
        double s = 0.0;
        for (int64_t i = b; i < n && i < b + 1024; i++) {
            const double u = q[4 * i + 2], v = q[4 * i + 3];
            s = s + 0.5 * (u * u + v * v);
        }
        total = b == 0 ? s : total + s;
    }
    return total;
}

typedef struct {
    float* q;       /* particles, 4 floats each */
    int64_t n, cap;
    float* rho;     /* (ny+4) x (nx+4) */
    float* phi;
    float* ex;      /* ny x nx */
    float* ey;
    int* count;     /* ny x nx */
} pic_rank;

static void pic_deposit(const orc_pic_params* p, pic_rank* R, int x0, int y0) {
    const int w = p->nx, h = p->ny, W = w + 4;
    memset(R->count, 0, sizeof(int) * (size_t)w * h);
    for (int64_t i = 0; i < R->n; i++) {
        const int lx = (int)R->q[4 * i] - x0, ly = (int)R->q[4 * i + 1] - y0;
        R->count[(size_t)ly * w + lx]++;
    }
    memset(R->rho, 0, sizeof(float) * (size_t)(h + 4) * W);
    for (int j = 0; j < h; j++)
        for (int i = 0; i < w; i++) R->rho[(size_t)(j + 2) * W + i + 2] = (float)R->count[(size_t)j * w + i];
}

/* The strip of rank A's interior that its neighbour in direction d needs,
 * copied into neighbour B's ring on the side facing A (B = A + d). */
static void pic_halo(const orc_pic_params* p, pic_rank* R) {
    const int P = p->px * p->py, w = p->nx, h = p->ny, W = w + 4;
    for (int d = 1; d <= 8; d++) {
        for (int b = 0; b < P; b++) {
            const int bx = b % p->px, by = b / p->px;
            /* receives, in step d, from the neighbour opposite d */
            const int ax = (bx - PIC_DX[d] + p->px) % p->px, ay = (by - PIC_DY[d] + p->py) % p->py;
            const pic_rank* A = &R[ay * p->px + ax];
            const int dx = PIC_DX[d], dy = PIC_DY[d];
            /* source rows/cols in A's interior (facing +d), target in B's ring (facing -d) */
            const int sx0 = dx < 0 ? 2 : dx > 0 ? w : 2, nxs = dx == 0 ? w : 2;
            const int sy0 = dy < 0 ? 2 : dy > 0 ? h : 2, nys = dy == 0 ? h : 2;
            const int tx0 = dx < 0 ? w + 2 : dx > 0 ? 0 : 2;
            const int ty0 = dy < 0 ? h + 2 : dy > 0 ? 0 : 2;
            for (int j = 0; j < nys; j++)
                for (int i = 0; i < nxs; i++)
                    R[b].rho[(size_t)(ty0 + j) * W + tx0 + i] = A->rho[(size_t)(sy0 + j) * W + sx0 + i];
        }
    }
}

int orc_pic_run(const orc_pic_params* p, double* energy, float* energy_bf16, int64_t* counts,
                uint64_t* part_hash, uint64_t* rho_hash, float* final_q_out) {
    const int P = p->px * p->py, w = p->nx, h = p->ny, W = w + 4;
    pic_rank* R = (pic_rank*)calloc((size_t)P, sizeof(pic_rank));
    if (!R) return -1;
    int64_t cap = p->np * 2 + 1024;
    for (int r = 0; r < P; r++) {
        R[r].cap = cap;
        R[r].q = (float*)malloc(sizeof(float) * 4 * (size_t)cap);
        R[r].rho = (float*)calloc((size_t)(h + 4) * W, sizeof(float));
        R[r].phi = (float*)calloc((size_t)(h + 4) * W, sizeof(float));
        R[r].ex = (float*)calloc((size_t)h * w, sizeof(float));
        R[r].ey = (float*)calloc((size_t)h * w, sizeof(float));
        R[r].count = (int*)calloc((size_t)h * w, sizeof(int));
        if (!R[r].q || !R[r].rho || !R[r].phi || !R[r].ex || !R[r].ey || !R[r].count) return -1;
        R[r].n = p->np;
        for (int64_t i = 0; i < p->np; i++) orc_pic_init_particle(p, r, i, R[r].q + 4 * i);
    }
    for (int r = 0; r < P; r++) pic_deposit(p, &R[r], (r % p->px) * w, (r / p->px) * h);
    pic_halo(p, R);
    for (int r = 0; r < P; r++) pic_field(R[r].rho, R[r].phi, R[r].ex, R[r].ey, w, h);
    float** out = (float**)calloc((size_t)P * 9, sizeof(float*));
    int64_t* nout = (int64_t*)calloc((size_t)P * 9, sizeof(int64_t));
    for (int c = 0; c < p->cycles; c++) {
        /* push, classify, stable partition */
        for (int r = 0; r < P; r++) {
            const int rx = r % p->px, ry = r / p->px, x0 = rx * w, y0 = ry * h;
            for (int d = 0; d < 9; d++) {
                free(out[r * 9 + d]);
                out[r * 9 + d] = (float*)malloc(sizeof(float) * 4 * (size_t)(R[r].n + 1));
                nout[r * 9 + d] = 0;
            }
            for (int64_t i = 0; i < R[r].n; i++) {
                float* q = R[r].q + 4 * i;
                pic_push(p, q, R[r].ex, R[r].ey, w, x0, y0);
                const int ox = (int)q[0] / w, oy = (int)q[1] / h;
                const int dx = ox == rx ? 0 : ox == (rx + 1) % p->px ? 1 : -1;
                const int dy = oy == ry ? 0 : oy == (ry + 1) % p->py ? 1 : -1;
                int d = 0;
                for (int k = 1; k <= 8; k++)
                    if (PIC_DX[k] == dx && PIC_DY[k] == dy) d = k;
                memcpy(out[r * 9 + d] + 4 * nout[r * 9 + d]++, q, sizeof(float) * 4);
            }
        }
        /* exchange: stays first, then step d's arrivals from the neighbour opposite d */
        for (int r = 0; r < P; r++) {
            const int rx = r % p->px, ry = r / p->px;
            int64_t n = 0;
            memcpy(R[r].q, out[r * 9], sizeof(float) * 4 * (size_t)nout[r * 9]);
            n = nout[r * 9];
            for (int d = 1; d <= 8; d++) {
                const int ax = (rx - PIC_DX[d] + p->px) % p->px, ay = (ry - PIC_DY[d] + p->py) % p->py;
                const int a = ay * p->px + ax;
                if (n + nout[a * 9 + d] > R[r].cap) return -2;
                memcpy(R[r].q + 4 * n, out[a * 9 + d], sizeof(float) * 4 * (size_t)nout[a * 9 + d]);
                n += nout[a * 9 + d];
            }
            R[r].n = n;
        }
        /* deposit, energy, halo, field */
        double e = 0.0;
        float e16 = 0.0f;
        for (int r = 0; r < P; r++) {
            pic_deposit(p, &R[r], (r % p->px) * w, (r / p->px) * h);
            const double er = orc_pic_energy(R[r].q, R[r].n);
            e = r == 0 ? er : er + e;
            const float d16 = orc_bf16_to_f32(orc_f32_to_bf16((float)er));
            e16 = r == 0 ? d16 : d16 + e16;
        }
        if (energy) energy[c] = e;
        if (energy_bf16) energy_bf16[c] = e16;
        pic_halo(p, R);
        for (int r = 0; r < P; r++) pic_field(R[r].rho, R[r].phi, R[r].ex, R[r].ey, w, h);
    }
    int64_t at = 0;
    for (int r = 0; r < P; r++) {
        if (final_q_out) memcpy(final_q_out + 4 * at, R[r].q, sizeof(float) * 4 * (size_t)R[r].n);
        at += R[r].n;
        if (counts) counts[r] = R[r].n;
        if (part_hash) part_hash[r] = pic_hash(R[r].q, sizeof(float) * 4 * (size_t)R[r].n);
        if (rho_hash) rho_hash[r] = pic_hash(R[r].rho, sizeof(float) * (size_t)(h + 4) * W);
        free(R[r].q); free(R[r].rho); free(R[r].phi); free(R[r].ex); free(R[r].ey); free(R[r].count);
    }
    for (int k = 0; k < P * 9; k++) free(out[k]);
    free(out); free(nout); free(R);
    return 0;
}

/* The same physics on the undecomposed (px*nx) x (py*ny) periodic grid: one
 * particle array in initial (rank, index) order, one density and field. */
int orc_pic_global(const orc_pic_params* p, double* energy, float* final_q_sorted_out, int64_t* n_out) {
    const int P = p->px * p->py, GW = p->px * p->nx, GH = p->py * p->ny, W = GW + 4;
    const int64_t n = (int64_t)P * p->np;
    float* q = (float*)malloc(sizeof(float) * 4 * (size_t)n);
    float* rho = (float*)calloc((size_t)(GH + 4) * W, sizeof(float));
    float* phi = (float*)calloc((size_t)(GH + 4) * W, sizeof(float));
    float* ex = (float*)calloc((size_t)GH * GW, sizeof(float));
    float* ey = (float*)calloc((size_t)GH * GW, sizeof(float));
    int* cnt = (int*)calloc((size_t)GH * GW, sizeof(int));
    if (!q || !rho || !phi || !ex || !ey || !cnt) return -1;
    for (int r = 0; r < P; r++)
        for (int64_t i = 0; i < p->np; i++) orc_pic_init_particle(p, r, i, q + 4 * ((int64_t)r * p->np + i));
    for (int c = -1; c < p->cycles; c++) {
        if (c >= 0)
            for (int64_t i = 0; i < n; i++) pic_push(p, q + 4 * i, ex, ey, GW, 0, 0);
        memset(cnt, 0, sizeof(int) * (size_t)GW * GH);
        for (int64_t i = 0; i < n; i++) cnt[(size_t)(int)q[4 * i + 1] * GW + (int)q[4 * i]]++;
        for (int j = 0; j < GH + 4; j++)
            for (int i = 0; i < GW + 4; i++) {
                const int gx = (i - 2 + GW) % GW, gy = (j - 2 + GH) % GH;
                rho[(size_t)j * W + i] = (float)cnt[(size_t)gy * GW + gx];
            }
        pic_field(rho, phi, ex, ey, GW, GH);
        if (c >= 0 && energy) {
            double s = 0.0;
            for (int64_t i = 0; i < n; i++) {
                const double u = q[4 * i + 2], v = q[4 * i + 3];
                s = s + 0.5 * (u * u + v * v);
            }
            energy[c] = s;
        }
    }
    if (final_q_sorted_out) memcpy(final_q_sorted_out, q, sizeof(float) * 4 * (size_t)n);
    if (n_out) *n_out = n;
    free(q); free(rho); free(phi); free(ex); free(ey); free(cnt);
    return 0;
}
