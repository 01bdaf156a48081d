// Payload movement.  Replaces the reference's single-thread byte loop
// (gpu_libs/gpu_libc/string.cu:185-192, used by cuda_mpi.cu:168,192,323,385)
// with cooperative copies: worker t of nt moves 16-byte vectors (dwordx4)
// strided by nt, several loads in flight before the matching stores.
#pragma once

#include "device.hpp"

namespace gpumpi {
namespace rt {

// The calling lane's view of a private (scratch) address kept in shared
// state (a request's buffer, say): on the GPU one private address names
// every lane's own copy, so it is the address itself.  The CPU simulation of
// the runtime (tests/hostbuild/simt) defines it to map the address into the
// calling lane's stack.
#ifndef GPUMPI_LANE_PRIVATE
#define GPUMPI_LANE_PRIVATE(p) (p)
#endif

__device__ __forceinline__ bool is_private_ptr(const void* p) {
    return __builtin_amdgcn_is_private((const __attribute__((address_space(0))) void*)p);
}
__device__ __forceinline__ bool is_shared_ptr(const void* p) {
    return __builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)p);
}
// Memory other workgroups can read or write (not scratch, not LDS).
__device__ __forceinline__ bool is_global_ptr(const void* p) {
    return !is_private_ptr(p) && !is_shared_ptr(p);
}

constexpr int kCopyUnroll = 8;

// Global-memory (HBM) pointers: loads and stores through them are global_*
// instructions, which count on vmcnt alone and cannot alias the stack.  (The
// host build of this header for the sanitizer checks predefines it empty.)
#ifndef GPUMPI_GLOBAL
#define GPUMPI_GLOBAL __attribute__((address_space(1)))
#endif
// LDS (local data share) pointers: atomics and loads through them are ds_*
// instructions rather than flat ones (host builds predefine it empty).
#ifndef GPUMPI_LDS
#define GPUMPI_LDS __attribute__((address_space(3)))
#endif
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16-byte vector, plain value type
typedef GPUMPI_GLOBAL u32x4 gvec4;

// Bytes [4*Q + r, 4*Q + r + 16) of the 32-byte concatenation lo:hi.
template <int Q>
__device__ __forceinline__ u32x4 realign16(const u32x4 lo, const u32x4 hi, uint32_t r) {
    const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    u32x4 o;
    o.x = __builtin_amdgcn_alignbyte(w[Q + 1], w[Q], r);
    o.y = __builtin_amdgcn_alignbyte(w[Q + 2], w[Q + 1], r);
    o.z = __builtin_amdgcn_alignbyte(w[Q + 3], w[Q + 2], r);
    o.w = __builtin_amdgcn_alignbyte(w[Q + 4], w[Q + 3], r);
    return o;
}

// d4[i] = s4[i] for i < nvec, worker t of nt taking i = t, t + nt, ...;
// kCopyUnroll loads in flight per lane before the matching stores (a wave
// moves 64 * 16 * kCopyUnroll bytes per round trip).  The values are named
// scalars, not an array: an array here was kept in scratch (dead 16-byte
// stack stores per iteration) and its stores forced a wait after every load.
template <typename DP, typename SP>
__device__ __forceinline__ void copy_aligned(DP d4, SP s4, uint64_t nvec, int t, int nt) {
    static_assert(kCopyUnroll == 8, "copy_aligned moves eight vectors per step");
    uint64_t i = (uint64_t)t;
    const uint64_t st = (uint64_t)nt;
    for (; i + 7 * st < nvec; i += 8 * st) {
        const u32x4 v0 = s4[i], v1 = s4[i + st], v2 = s4[i + 2 * st], v3 = s4[i + 3 * st];
        const u32x4 v4 = s4[i + 4 * st], v5 = s4[i + 5 * st], v6 = s4[i + 6 * st], v7 = s4[i + 7 * st];
        d4[i] = v0;
        d4[i + st] = v1;
        d4[i + 2 * st] = v2;
        d4[i + 3 * st] = v3;
        d4[i + 4 * st] = v4;
        d4[i + 5 * st] = v5;
        d4[i + 6 * st] = v6;
        d4[i + 7 * st] = v7;
    }
    for (; i + 3 * st < nvec; i += 4 * st) {
        const u32x4 v0 = s4[i], v1 = s4[i + st], v2 = s4[i + 2 * st], v3 = s4[i + 3 * st];
        d4[i] = v0;
        d4[i + st] = v1;
        d4[i + 2 * st] = v2;
        d4[i + 3 * st] = v3;
    }
    for (; i < nvec; i += st) d4[i] = s4[i];
}

// d4[i] = source bytes [16 i + 4 Q + r, +16) for i < nvec, reading only the
// 16-byte aligned source blocks a4[i] and a4[i + 1]: every block read holds
// at least one byte of the source range, so no read leaves its page.
template <int Q>
__device__ inline void copy_realigned(gvec4* d4, const gvec4* a4, uint64_t nvec, uint32_t r, int t, int nt) {
    uint64_t i = (uint64_t)t;
    const uint64_t st = (uint64_t)nt;
    for (; i + 7 * st < nvec; i += 8 * st) {
        u32x4 l[8], h[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            l[u] = a4[i + u * st];
            h[u] = a4[i + u * st + 1];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) d4[i + u * st] = realign16<Q>(l[u], h[u], r);
    }
    for (; i + 3 * st < nvec; i += 4 * st) {
        const u32x4 l0 = a4[i], h0 = a4[i + 1];
        const u32x4 l1 = a4[i + st], h1 = a4[i + st + 1];
        const u32x4 l2 = a4[i + 2 * st], h2 = a4[i + 2 * st + 1];
        const u32x4 l3 = a4[i + 3 * st], h3 = a4[i + 3 * st + 1];
        d4[i] = realign16<Q>(l0, h0, r);
        d4[i + st] = realign16<Q>(l1, h1, r);
        d4[i + 2 * st] = realign16<Q>(l2, h2, r);
        d4[i + 3 * st] = realign16<Q>(l3, h3, r);
    }
    for (; i < nvec; i += st) d4[i] = realign16<Q>(a4[i], a4[i + 1], r);
}

// Worker t of nt copies its share of [src, src+n) to dst.  Every worker
// passes the same arguments; nt >= 16.  Buffers may overlap only if
// dst == src.  The destination is aligned to 16 bytes by a short head, then
// moved in 16-byte vectors (dwordx4 loads and stores, several in flight per
// lane); a source at another alignment is read as aligned vectors and
// shifted into place.  Global (HBM) buffers take global_* instructions;
// a private or LDS buffer takes the generic (flat) path.
__device__ inline void copy_part_lanes(void* dstv, const void* srcv, uint64_t n, int t, int nt) {
    uint8_t* dst = (uint8_t*)dstv;
    const uint8_t* src = (const uint8_t*)srcv;
    if (n == 0 || dst == src) return;
    uint64_t head = (16 - ((uintptr_t)dst & 15)) & 15;
    if (head > n) head = n;
    if ((uint64_t)t < head) dst[t] = src[t];
    const uint64_t nvec = (n - head) >> 4;
    uint8_t* d2 = dst + head;
    const uint8_t* s2 = src + head;
    const uint32_t m = (uint32_t)((uintptr_t)s2 & 15);
    const bool global = is_global_ptr(src) && is_global_ptr(dst);
    if (m == 0 && global) {
        copy_aligned((gvec4*)d2, (const gvec4*)s2, nvec, t, nt);
    } else if (m == 0) {
        copy_aligned((u32x4*)d2, (const u32x4*)s2, nvec, t, nt);
    } else if (global) {
        gvec4* d4 = (gvec4*)d2;
        const gvec4* a4 = (const gvec4*)(s2 - m);
        switch (m >> 2) {
        case 0: copy_realigned<0>(d4, a4, nvec, m & 3, t, nt); break;
        case 1: copy_realigned<1>(d4, a4, nvec, m & 3, t, nt); break;
        case 2: copy_realigned<2>(d4, a4, nvec, m & 3, t, nt); break;
        default: copy_realigned<3>(d4, a4, nvec, m & 3, t, nt); break;
        }
    } else {
        // Private or LDS buffer at another alignment: plain bytes (whole
        // aligned blocks are read only from global memory, where a block
        // never leaves its page).
        for (uint64_t i = head + (uint64_t)t; i < n; i += (uint64_t)nt) dst[i] = src[i];
        return;
    }
    const uint64_t done = head + (nvec << 4);
    if ((uint64_t)t < n - done) dst[done + t] = src[done + t];
}
// The wave's lanes may read what the others copied as soon as this returns.
__device__ inline void copy_part(void* dstv, const void* srcv, uint64_t n, int t, int nt) {
    copy_part_lanes(dstv, srcv, n, t, nt);
    wave_sync();
}

// Every lane copies all n bytes: used when dst is the caller's private
// (scratch) memory, of which each lane of wave 0 holds its own copy.
__device__ inline void copy_all_lanes(void* dstv, const void* srcv, uint64_t n) {
    uint8_t* dst = (uint8_t*)dstv;
    const uint8_t* src = (const uint8_t*)srcv;
    if ((((uintptr_t)dst | (uintptr_t)src) & 7) == 0) {
        uint64_t nw = n >> 3;
        for (uint64_t i = 0; i < nw; i++) ((uint64_t*)dst)[i] = ((const uint64_t*)src)[i];
        for (uint64_t i = nw << 3; i < n; i++) dst[i] = src[i];
    } else if ((((uintptr_t)dst | (uintptr_t)src) & 3) == 0) {
        uint64_t nw = n >> 2;
        for (uint64_t i = 0; i < nw; i++) ((uint32_t*)dst)[i] = ((const uint32_t*)src)[i];
        for (uint64_t i = nw << 2; i < n; i++) dst[i] = src[i];
    } else {
        for (uint64_t i = 0; i < n; i++) dst[i] = src[i];
    }
}

}  // namespace rt
}  // namespace gpumpi
