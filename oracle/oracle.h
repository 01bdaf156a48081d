/*
 * CPU oracle for the device-resident MPI hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it restates, in
 * plain C, the arithmetic and data movement the reference MPI runtime
 * (lilyyangyn/dphpc-gpu_mpi) performs on the hot path, so the HIP runtime
 * can be compared against it on the same inputs.  Only tests/, the smoke
 * entry point and bench.py's CPU-baseline leg may load it.  The product
 * path never calls into this file.
 *
 * Each function cites the reference file:line it follows.  How each one is
 * pinned (golden vectors, closed forms, host MPI) is listed next to it; a
 * function marked "parity unpinned" has no independent check.
 */
#ifndef GPUMPI_ORACLE_H
#define GPUMPI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- datatypes (gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33) ------------- */
/* Enum order and element sizes; MPI_LONG_DOUBLE is an 8-byte double and the
 * long-double complex a 16-byte double complex, as in the reference. */
int orc_type_size(int dtype);
/* 1 if (dtype, op) is a legal reduction pair (operators.cu:153-200). */
int orc_op_allowed(int dtype, int op);

/* ---- reductions (gpu_libs/gpu_mpi/mpi.cu:188-234, operators.cu:73-218) --
 * Generic Reduce: acc = x_0, then acc[i] = op(x_r[i], acc[i]) for r=1..P-1
 * in rank order.  `inputs` holds P contiguous vectors of `count` elements.
 * Returns 0, or -1 if the pair is illegal.
 * Pinned by: closed forms (sum of (r+1)(i+1)), the int16 PROD wraparound
 * known-answer test (tests/collectives_test.cu:6-54), and MPICH output for
 * every legal (type, op) pair on integer-valued inputs (tests/golden). */
int orc_reduce_fold(int dtype, int op, int nranks, size_t count,
                    const void* inputs, void* out);

/* ---- pi by the midpoint rule (toolchain_tests/pi/cpi.c:41-56) ----------
 * Partial sum of rank `rank` of `nranks` for n intervals (cyclic split
 * i = rank+1, rank+1+nranks, ...), then the rank-order fold of the partials.
 * Pinned by: closed form pi (cpi.c:23) and the reference cpi.c itself run
 * under host MPI. */
double orc_pi_partial(int n, int rank, int nranks);
double orc_pi_fold(int n, int nranks, double* partials_out);

/* ---- seeded payloads ----------------------------------------------------
 * Counter-based 64-bit words: word i of rank r under seed s.  Bytes are the
 * little-endian bytes of consecutive words.  Pinned by: splitmix64's
 * published test vector (seed 1234567). */
uint64_t orc_mix64(uint64_t x);
uint64_t orc_payload_word(uint64_t seed, uint64_t rank, uint64_t i);
void orc_payload_bytes(uint64_t seed, uint64_t rank, size_t nbytes, uint8_t* out);
uint64_t orc_splitmix64_next(uint64_t* state);

/* ---- NPB IS (toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c) ----------
 * randlc (is.c:408-476), find_my_seed (:489-533), create_seq (:535-553),
 * rank (:631-930), full_verify (:561-625), main loop (:1060-1137).
 * Pinned by: NPB's own known-answer arrays (is.c:254-282, :869-913) and the
 * full-verify pass count 5*MAX_ITERATIONS + comm_size (is.c:1137). */
double orc_randlc(double* x, double a);
double orc_find_my_seed(int kn, int np, long nn, double s, double a);

typedef struct orc_is_result {
    int passed_verification;  /* summed over ranks, as in is.c:1113-1137 */
    int full_verify_errors;   /* keys out of order, summed over ranks */
    int64_t total_keys_after; /* sum of total_local_keys over ranks */
    uint64_t sorted_checksum; /* sum over ranks of sum_i mix64(key_i + i*2^32) */
    double timed_seconds;     /* wall time of the 10 timed iterations (one host thread) */
} orc_is_result;

/* Runs the whole IS benchmark (one untimed + 10 iterations + full_verify)
 * for `cls` in {'S','W','A','B'} on `nranks` simulated ranks.  If
 * `keys_out` is non-NULL it receives the globally sorted keys (TOTAL_KEYS).
 * If bucket_totals_out is non-NULL it receives bucket_size_totals of the
 * last iteration (NUM_BUCKETS + 5 ints). */
int orc_is_run(char cls, int nranks, orc_is_result* res, int32_t* keys_out,
               int32_t* bucket_totals_out);
/* Keys of rank `rank` right after create_seq (num_keys of them). */
int orc_is_keys(char cls, int nranks, int rank, int32_t* out);

/* ---- compressed wire codecs (BF16, OCP FP8 e4m3 with per-block scale) --
 * Written from the published formats: BF16 = top half of binary32 with
 * round-to-nearest-even; FP8 e4m3 (OCP "fn": bias 7, max 448, no inf,
 * S.1111.111 = NaN), saturating, with an FP32 scale amax/448 per block.
 * Pinned by: torch's bfloat16 / float8_e4m3fn casts on CPU. */
uint16_t orc_f32_to_bf16(float x);
float orc_bf16_to_f32(uint16_t h);
uint8_t orc_f32_to_e4m3(float x);
float orc_e4m3_to_f32(uint8_t q);
/* Quantise n floats in blocks of `block`: q[n] codes and scales[ceil(n/block)]. */
void orc_e4m3_block_encode(const float* x, size_t n, int block, uint8_t* q, float* scales);
void orc_e4m3_block_decode(const uint8_t* q, const float* scales, size_t n, int block, float* y);

/* ---- communicator split (gpu_libs/gpu_mpi/communicator.cu:147-213) ----
 * For each rank: new rank = position in its colour's member list ordered by
 * (key, parent rank); size = members of that colour.  Pinned by the
 * known-answer table of tests/communicator_test.cu:93-152. */
void orc_comm_split(int nranks, const int* colors, const int* keys,
                    int* new_rank, int* new_size);

/* ---- 2-D particle-in-cell cycle (the C5 workload) ---------------------
 * This build's own workload (the reference has no spatially decomposed PIC;
 * its sputniPIC, toolchain_tests/sputnipic/src/mpi_comm.c:21-307, scatters
 * particles).  Ranks form a px x py grid; rank r = ry*px + rx owns cells
 * [rx*nx, rx*nx+nx) x [ry*ny, ry*ny+ny) of a periodic domain, unit cells.
 * Particles are (x, y, u, v) in FP32, positions global.  All FP arithmetic
 * is written out in the order below and built without contraction.
 *   init   draw k of particle i of rank r: mix64(seed + r*0xD1B54A32D192ED03
 *          + (32i+k+1)*0x9E3779B97F4A7C15), unit = (w>>40)*2^-24;
 *          x = x0 + nx*unit(k=0) (x0 if that rounds to x0+nx), y likewise
 *          (k=1); u = drift + vth*(sum_k=2..13 unit - 6), v = vth*(sum_k=14..25
 *          unit - 6) (Irwin-Hall normals, summed in k order).
 *   cycle  push: u += qm*Ex(cell), v += qm*Ey(cell), x += u, y += v, wrapped
 *          once into [0, L); classify by the owner of (int)x, (int)y: stay
 *          or direction d = 1..8 of (dx, dy) in row order (-1,-1) .. (1,1);
 *          stable partition; exchange: the rank keeps its stays in order,
 *          then for d = 1..8 appends what the neighbour opposite d sent in
 *          direction d; deposit NGP counts; energy = sum over particles of
 *          0.5*(u*u + v*v) in FP64, 1024-particle blocks summed in order per
 *          rank, then folded in rank order (MPI_Allreduce), and the same per
 *          rank in FP32 through the BF16 wire (fold of dequantised values);
 *          rho = counts with a two-cell ring filled, for d = 1..8, from the
 *          neighbour opposite d; phi = ((((4c + w) + e) + s) + n)/8 on the
 *          interior and one ring; E = ((phi[-1] - phi[+1])/2, (phi[-W] -
 *          phi[+W])/2) on the interior.  The fields are also built once
 *          before the first cycle.
 * Pinned by: orc_pic_global, the same physics on the undecomposed grid
 * (identical particle states as a multiset, identical densities, energies
 * within FP64 summation-order rounding); particle-count conservation; and
 * with qm = 0 the ballistic closed form (constant energy, per-particle
 * velocities unchanged). */
typedef struct orc_pic_params {
    int px, py, nx, ny;
    int64_t np;        /* initial particles per rank */
    int cycles;
    uint64_t seed;
    float vth, drift, qm;
} orc_pic_params;
void orc_pic_init_particle(const orc_pic_params* p, int rank, int64_t i, float* xyuv);
double orc_pic_energy(const float* q, int64_t n);
/* energy[cycles] (FP64 fold), energy_bf16[cycles], and per rank the final
 * particle count, an FNV-1a hash of its particle array and of its rho
 * (with ring); final_q_out (4 floats per particle, ranks in order) if
 * non-NULL.  Returns 0, or <0 on allocation failure. */
int orc_pic_run(const orc_pic_params* p, double* energy, float* energy_bf16, int64_t* counts,
                uint64_t* part_hash, uint64_t* rho_hash, float* final_q_out);
int orc_pic_global(const orc_pic_params* p, double* energy, float* final_q_out, int64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif
