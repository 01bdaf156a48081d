// Payload movement.  Replaces the reference's single-thread byte loop
// (gpu_libs/gpu_libc/string.cu:185-192, used by cuda_mpi.cu:168,192,323,385)
// with cooperative copies: worker t of nt moves 16-byte vectors (dwordx4)
// strided by nt, several loads in flight before the matching stores.
#pragma once

#include "device.hpp"

namespace gpumpi {
namespace rt {

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

constexpr int kCopyUnroll = 4;

// Worker t of nt copies its share of [src, src+n) to dst.  Every worker
// passes the same arguments.  Buffers may overlap only if dst == src.
__device__ inline void copy_part(void* dstv, const void* srcv, uint64_t n, int t, int nt) {
    uint8_t* dst = (uint8_t*)dstv;
    const uint8_t* src = (const uint8_t*)srcv;
    if (n == 0 || dst == src) return;
    const uintptr_t d = (uintptr_t)dst, s = (uintptr_t)src;
    if (((d ^ s) & 15) == 0) {
        uint64_t head = (16 - (d & 15)) & 15;
        if (head > n) head = n;
        if ((uint64_t)t < head) dst[t] = src[t];
        const uint64_t nvec = (n - head) >> 4;
        const uint4* s4 = (const uint4*)(src + head);
        uint4* d4 = (uint4*)(dst + head);
        uint64_t i = (uint64_t)t;
        const uint64_t st = (uint64_t)nt;
        for (; i + (kCopyUnroll - 1) * st < nvec; i += kCopyUnroll * st) {
            uint4 v[kCopyUnroll];
#pragma unroll
            for (int u = 0; u < kCopyUnroll; u++) v[u] = s4[i + u * st];
#pragma unroll
            for (int u = 0; u < kCopyUnroll; u++) d4[i + u * st] = v[u];
        }
        for (; i < nvec; i += st) d4[i] = s4[i];
        const uint64_t done = head + (nvec << 4);
        const uint64_t tail = n - done;
        if ((uint64_t)t < tail) dst[done + t] = src[done + t];
        return;
    }
    if (((d ^ s) & 3) == 0) {
        uint64_t head = (4 - (d & 3)) & 3;
        if (head > n) head = n;
        if ((uint64_t)t < head) dst[t] = src[t];
        const uint64_t nw = (n - head) >> 2;
        const uint32_t* s1 = (const uint32_t*)(src + head);
        uint32_t* d1 = (uint32_t*)(dst + head);
        for (uint64_t i = (uint64_t)t; i < nw; i += (uint64_t)nt) d1[i] = s1[i];
        const uint64_t done = head + (nw << 2);
        if ((uint64_t)t < n - done) dst[done + t] = src[done + t];
        return;
    }
    for (uint64_t i = (uint64_t)t; i < n; i += (uint64_t)nt) dst[i] = src[i];
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
