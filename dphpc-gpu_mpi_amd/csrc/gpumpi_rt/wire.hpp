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

// Wire image of n floats: codes for n rounded up to 16 elements (BF16: 2 B,
// FP8: 1 B each), then, for FP8, one FP32 scale per 256-element block.  The
// fold reads codes in whole 16-element groups; codes past n are never
// written out.
constexpr int kWireGroup = 16;
__device__ __forceinline__ uint64_t wire_n16(uint64_t n) { return (n + 15) & ~(uint64_t)15; }
__device__ __forceinline__ uint64_t wire_bytes(int fmt, uint64_t n) {
    if (fmt == WIRE_BF16) return 2 * wire_n16(n);
    const uint64_t nblk = (n + kFp8Block - 1) / kFp8Block;
    return wire_n16(n) + 4 * nblk;
}
__device__ __forceinline__ float* wire_scales(uint8_t* w, uint64_t n) { return (float*)(w + wire_n16(n)); }

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m));
    return v;
}

// CMD_USER bodies are lambdas; the quantiser and the folding pass below are
// run on every wave through par_run.
struct QuantArgs {
    const float* x;
    uint8_t* wire;
    uint64_t n;
    int fmt;
};
// BF16: worker t of nt converts 8-element groups (two 16-byte loads, one
// 16-byte store).  FP8: wave w of the rank takes 256-element blocks, four
// elements per lane: a butterfly gives the block's amax, then every lane
// stores its four codes as one word and lane 0 the block's scale.
__device__ inline void quantise_part(const QuantArgs& A, int t, int nt) {
    const uint64_t n = A.n;
    const bool vec = ((uintptr_t)A.x & 15) == 0;
    if (A.fmt == WIRE_BF16) {
        uint16_t* w = (uint16_t*)A.wire;
        const uint64_t ng = n / 8;
        if (vec) {
            for (uint64_t g = (uint64_t)t; g < ng; g += (uint64_t)nt) {
                const float4 a = ((const float4*)A.x)[2 * g], b = ((const float4*)A.x)[2 * g + 1];
                uint4 o;
                o.x = (uint32_t)f32_to_bf16(a.x) | ((uint32_t)f32_to_bf16(a.y) << 16);
                o.y = (uint32_t)f32_to_bf16(a.z) | ((uint32_t)f32_to_bf16(a.w) << 16);
                o.z = (uint32_t)f32_to_bf16(b.x) | ((uint32_t)f32_to_bf16(b.y) << 16);
                o.w = (uint32_t)f32_to_bf16(b.z) | ((uint32_t)f32_to_bf16(b.w) << 16);
                ((uint4*)w)[g] = o;
            }
        }
        for (uint64_t i = (vec ? 8 * ng : 0) + (uint64_t)t; i < n; i += (uint64_t)nt) w[i] = f32_to_bf16(A.x[i]);
        return;
    }
    const int lane = t & 63, wave = t >> 6, nwaves = nt >> 6;
    const uint64_t nblk = (n + kFp8Block - 1) / kFp8Block;
    float* scales = wire_scales(A.wire, n);
    for (uint64_t b = (uint64_t)wave; b < nblk; b += (uint64_t)nwaves) {
        const uint64_t e0 = b * kFp8Block + 4 * (uint64_t)lane;
        float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const bool full = e0 + 4 <= n;
        if (full && vec) {
            const float4 q = *(const float4*)(A.x + e0);
            v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) v[k] = A.x[e0 + k];
        }
        float amax = fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3])));
        amax = wave_max(amax);
        const float scale = amax > 0.0f ? amax / 448.0f : 1.0f;
        const float inv = amax > 0.0f ? 448.0f / amax : 1.0f;
        uint32_t c[4];
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] = f32_to_e4m3(v[k] * inv);
        if (full) {
            *(uint32_t*)(A.wire + e0) = c[0] | (c[1] << 8) | (c[2] << 16) | (c[3] << 24);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) A.wire[e0 + k] = (uint8_t)c[k];
        }
        if (lane == 0) scales[b] = scale;
    }
}

struct WireFoldArgs {
    const uint64_t* mem;  // members_tab(): [2r] wire image, [2r+1] FP32 output
    int nmembers;
    int fmt;
    uint64_t n;  // elements of the whole vector
    uint64_t g0, g1;  // this member's 16-element groups
    int vec_out;      // every member's output is 16-byte aligned
};
__device__ __forceinline__ float wire_load(int fmt, const uint8_t* w, uint64_t n, uint64_t e) {
    if (fmt == WIRE_BF16) return bf16_to_f32(((const uint16_t*)w)[e]);
    const float* scales = (const float*)(w + wire_n16(n));
    return e4m3_to_f32(w[e]) * scales[e / kFp8Block];
}
// The 16 values of group g of one member's wire image.
__device__ __forceinline__ void wire_group(int fmt, const uint8_t* w, uint64_t n, uint64_t g, float* v) {
    if (fmt == WIRE_BF16) {
        const uint4 a = ((const uint4*)w)[2 * g], b = ((const uint4*)w)[2 * g + 1];
        const uint32_t u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            v[2 * k] = __uint_as_float(u[k] << 16);
            v[2 * k + 1] = __uint_as_float(u[k] & 0xffff0000u);
        }
    } else {
        const uint4 a = ((const uint4*)w)[g];
        const float s = ((const float*)(w + wire_n16(n)))[(g * kWireGroup) / kFp8Block];
        const uint32_t u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = e4m3_to_f32((uint8_t)(u[k >> 2] >> (8 * (k & 3)))) * s;
    }
}
// Rank-order FP32 fold of the dequantised members over this member's
// groups: worker t of nt takes whole 16-element groups, B members' loads in
// flight, and writes the 16 sums to every member's output.  The last,
// partial group goes element by element.
__device__ inline void wire_fold_part(const WireFoldArgs& A, int t, int nt) {
    // One member per batch keeps this body (called through the helper
    // waves' indirect call) within the caller-saved registers, so it needs
    // no stack frame: the kernel's fixed private segment is sized from its
    // direct calls only, and a frame here would lie beyond it.
    constexpr int B = 1;
    const int P = A.nmembers;
    const uint64_t full_groups = A.n / kWireGroup;
    for (uint64_t g = A.g0 + (uint64_t)t; g < A.g1; g += (uint64_t)nt) {
        if (g < full_groups) {
            float acc[16];
            wire_group(A.fmt, (const uint8_t*)(uintptr_t)A.mem[0], A.n, g, acc);
            for (int rb = 1; rb < P; rb += B) {
                float x[B][16];
#pragma unroll
                for (int k = 0; k < B; k++)
                    if (rb + k < P) wire_group(A.fmt, (const uint8_t*)(uintptr_t)A.mem[2 * (rb + k)], A.n, g, x[k]);
#pragma unroll
                for (int k = 0; k < B; k++)
                    if (rb + k < P) {
#pragma unroll
                        for (int j = 0; j < 16; j++) acc[j] = x[k][j] + acc[j];
                    }
            }
            for (int r = 0; r < P; r++) {
                float* o = (float*)(uintptr_t)A.mem[2 * r + 1] + g * kWireGroup;
                if (A.vec_out) {
                    ((float4*)o)[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
                    ((float4*)o)[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
                    ((float4*)o)[2] = make_float4(acc[8], acc[9], acc[10], acc[11]);
                    ((float4*)o)[3] = make_float4(acc[12], acc[13], acc[14], acc[15]);
                } else {
#pragma unroll
                    for (int j = 0; j < 16; j++) o[j] = acc[j];
                }
            }
        } else {
            for (uint64_t e = g * kWireGroup; e < A.n; e++) {
                float acc = wire_load(A.fmt, (const uint8_t*)(uintptr_t)A.mem[0], A.n, e);
                for (int r = 1; r < P; r++)
                    acc = wire_load(A.fmt, (const uint8_t*)(uintptr_t)A.mem[2 * r], A.n, e) + acc;
                for (int r = 0; r < P; r++) ((float*)(uintptr_t)A.mem[2 * r + 1])[e] = acc;
            }
        }
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
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    const uint64_t G = (n + kWireGroup - 1) / kWireGroup;
    const uint64_t g0 = G * (uint64_t)cc.me / (uint64_t)cc.size, g1 = G * (uint64_t)(cc.me + 1) / (uint64_t)cc.size;
    if (g1 > g0) {
        const int vec_out = all_aligned16(cc, SW_RBUF, SW_RBUF, -1) ? 1 : 0;
        load_members(cc, SW_SBUF, SW_RBUF);
        WireFoldArgs f{members_tab(), cc.size, fmt, n, g0, g1, vec_out};
        par_run([=] __device__(int t, int nt) { wire_fold_part(f, t, nt); },
                (g1 - g0) * kWireGroup * (uint64_t)(cc.size) * 6);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
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
