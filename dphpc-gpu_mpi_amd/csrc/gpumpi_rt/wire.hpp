// Compressed-wire float Allreduce (opt-in; MPI_FLOAT + MPI_SUM only).
//
// Every member quantises its FP32 input once into a wire buffer on its rank
// heap -- BF16 (round to nearest even) or OCP FP8 e4m3 with one FP32 scale
// (amax/448) per 256-element block -- then the members fold the wire values
// in rank order with FP32 accumulation, exactly like the exact path
// (acc = x_0; acc = x_r + acc), and write FP32 results to every member.
// Integer and byte datatypes are refused: they never travel compressed.
#pragma once

#include "coll.hpp"

namespace gpumpi {
namespace rt {

enum WireFormat { WIRE_EXACT = 0, WIRE_BF16 = 1, WIRE_FP8_E4M3 = 2 };
constexpr int kFp8Block = 256;

__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// OCP e4m3 ("fn": bias 7, max 448, S.1111.111 = NaN, no inf), round to
// nearest even, saturating.
__device__ __forceinline__ uint8_t f32_to_e4m3(float x) {
    const uint32_t u = __float_as_uint(x);
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return sign | 0x7f;
    const float af = __uint_as_float(a);
    if (af >= 448.0f) return sign | 0x7e;
    if (af < 0.015625f) return sign | (uint8_t)rintf(af * 512.0f);  // subnormal grid 2^-9
    int e = (int)(a >> 23) - 127;
    const uint32_t m = a & 0x7fffffu;
    uint32_t r = m >> 20;
    const uint32_t rem = m & 0xfffffu;
    if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) r++;
    if (r == 8) {
        r = 0;
        e++;
    }
    uint32_t code = ((uint32_t)(e + 7) << 3) | r;
    if (code > 0x7e) code = 0x7e;
    return sign | (uint8_t)code;
}
__device__ __forceinline__ float e4m3_to_f32(uint8_t q) {
    const uint32_t e = (q >> 3) & 15u, m = q & 7u;
    float v;
    if (e == 15 && m == 7) v = __uint_as_float(0x7fc00000u);
    else if (e == 0) v = (float)m * 0.001953125f;                           // m * 2^-9
    else v = __uint_as_float(((e + 120u) << 23) | (m << 20));               // (1 + m/8) * 2^(e-7)
    return (q & 0x80) ? -v : v;
}

// Bytes of the wire image of n floats.
__device__ __forceinline__ uint64_t wire_bytes(int fmt, uint64_t n) {
    if (fmt == WIRE_BF16) return 2 * n;
    const uint64_t nblk = (n + kFp8Block - 1) / kFp8Block;
    return ((n + 15) & ~(uint64_t)15) + 4 * nblk;
}

// CMD_USER bodies are lambdas; the quantiser and the folding pass below are
// run on every wave through wg_run.
struct QuantArgs {
    const float* x;
    uint8_t* wire;
    uint64_t n;
    int fmt;
};
__device__ inline void quantise_part(const QuantArgs& A, int t, int nt) {
    if (A.fmt == WIRE_BF16) {
        uint16_t* w = (uint16_t*)A.wire;
        for (uint64_t i = (uint64_t)t; i < A.n; i += (uint64_t)nt) w[i] = f32_to_bf16(A.x[i]);
        return;
    }
    // FP8: thread t owns whole 256-element blocks.
    const uint64_t nblk = (A.n + kFp8Block - 1) / kFp8Block;
    float* scales = (float*)(A.wire + ((A.n + 15) & ~(uint64_t)15));
    for (uint64_t b = (uint64_t)t; b < nblk; b += (uint64_t)nt) {
        const uint64_t lo = b * kFp8Block, hi = min(A.n, lo + kFp8Block);
        float amax = 0.0f;
        for (uint64_t i = lo; i < hi; i++) amax = fmaxf(amax, fabsf(A.x[i]));
        const float scale = amax > 0.0f ? amax / 448.0f : 1.0f;
        const float inv = amax > 0.0f ? 448.0f / amax : 1.0f;
        scales[b] = scale;
        for (uint64_t i = lo; i < hi; i++) A.wire[i] = f32_to_e4m3(A.x[i] * inv);
    }
}

struct WireFoldArgs {
    const uint64_t* mem;  // members_lds(): [2r] wire image, [2r+1] FP32 output
    int nmembers;
    int fmt;
    uint64_t n;  // elements of the whole vector
    uint64_t lo, hi;
};
__device__ __forceinline__ float wire_load(int fmt, const uint8_t* w, uint64_t n, uint64_t e) {
    if (fmt == WIRE_BF16) return bf16_to_f32(((const uint16_t*)w)[e]);
    const float* scales = (const float*)(w + ((n + 15) & ~(uint64_t)15));
    return e4m3_to_f32(w[e]) * scales[e / kFp8Block];
}
// Rank-order FP32 fold of the dequantised members, kFoldBatch members'
// loads in flight at a time.
__device__ inline void wire_fold_part(const WireFoldArgs& A, int t, int nt) {
    constexpr int B = kFoldBatch;
    const int P = A.nmembers;
    for (uint64_t e = A.lo + (uint64_t)t; e < A.hi; e += (uint64_t)nt) {
        float acc = wire_load(A.fmt, (const uint8_t*)(uintptr_t)A.mem[0], A.n, e);
        for (int rb = 1; rb < P; rb += B) {
            float x[B];
#pragma unroll
            for (int k = 0; k < B; k++)
                if (rb + k < P) x[k] = wire_load(A.fmt, (const uint8_t*)(uintptr_t)A.mem[2 * (rb + k)], A.n, e);
#pragma unroll
            for (int k = 0; k < B; k++)
                if (rb + k < P) acc = x[k] + acc;
        }
        for (int r = 0; r < P; r++) ((float*)(uintptr_t)A.mem[2 * r + 1])[e] = acc;
    }
}

__device__ inline int wire_allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype dt, MPI_Op op,
                                     MPI_Comm comm, int fmt) {
    if (fmt == WIRE_EXACT) return reduce_common(sendbuf, recvbuf, count, dt, op, -1, comm);
    if (dt != MPI_FLOAT || op != MPI_SUM)
        return GPUMPI_FAIL(MPI_ERR_TYPE, "compressed Allreduce supports MPI_FLOAT with MPI_SUM only");
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    const void* src = sendbuf == MPI_IN_PLACE ? recvbuf : sendbuf;
    if (!is_global_ptr(src) || !is_global_ptr(recvbuf))
        return GPUMPI_FAIL(MPI_ERR_BUFFER, "compressed Allreduce needs global buffers");
    const uint64_t n = (uint64_t)count;
    const uint64_t wb = wire_bytes(fmt, n);
    uint8_t* wire = (uint8_t*)rank_malloc(wb);
    if (!wire) return MPI_ERR_NO_MEM;
    {
        QuantArgs q{(const float*)src, wire, n, fmt};
        par_run([=] __device__(int t, int nt) { quantise_part(q, t, nt); }, 4 * n + wb);
    }
    publish(cc, SW_SBUF, (uint64_t)(uintptr_t)wire);
    publish(cc, SW_RBUF, (uint64_t)(uintptr_t)recvbuf);
    comm_barrier(cc.c, cc.size);
    const uint64_t g0 = n * (uint64_t)cc.me / (uint64_t)cc.size, g1 = n * (uint64_t)(cc.me + 1) / (uint64_t)cc.size;
    if (g1 > g0) {
        load_members(cc, SW_SBUF, SW_RBUF);
        WireFoldArgs f{members_lds(), cc.size, fmt, n, g0, g1};
        par_run([=] __device__(int t, int nt) { wire_fold_part(f, t, nt); },
                           (g1 - g0) * (uint64_t)(cc.size) * 6);
    }
    comm_barrier(cc.c, cc.size);
    rank_free(wire, wb);
    return MPI_SUCCESS;
}

}  // namespace rt
}  // namespace gpumpi

// Opt-in compressed Allreduce (extension; no counterpart in the reference):
// wire = 0 exact (MPI_Allreduce), 1 BF16, 2 FP8 e4m3 with per-256 scales.
__device__ inline int GPUMPI_Allreduce_wire(const void* sendbuf, void* recvbuf, int count, MPI_Datatype dt,
                                            MPI_Op op, MPI_Comm comm, int wire) {
    return gpumpi::rt::wire_allreduce(sendbuf, recvbuf, count, dt, op, comm, wire);
}
