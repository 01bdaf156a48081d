// Host (CPU) build of the runtime headers, for tests only.
//
// The device runtime is written for 64-lane wavefronts.  This header lets a
// host compiler build it so that lane-local pieces (copies, segment
// addressing, reduction folds) can be run one lane at a time under
// AddressSanitizer: `gpumpi_host_lane` selects the lane, wave-wide
// operations (ballot, shuffles, readfirstlane, barriers) are not available
// and abort if reached.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ thread_local
#define GPUMPI_GLOBAL
#define GPUMPI_LDS
#define gpumpi_lds_base (*gpumpi_host_lds)

extern thread_local int gpumpi_host_lane;

[[noreturn]] inline void gpumpi_host_unsupported(const char* what) {
    std::fprintf(stderr, "host build: wave-wide operation %s reached\n", what);
    std::abort();
}

struct gpumpi_host_idx {
    unsigned x, y, z;
};
extern thread_local gpumpi_host_idx threadIdx, blockIdx, blockDim, gridDim;

struct alignas(16) uint4 {
    unsigned x, y, z, w;
};
struct alignas(16) float4 {
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }

inline unsigned __lane_id() { return (unsigned)gpumpi_host_lane; }
inline unsigned long long __ballot(int) { gpumpi_host_unsupported("ballot"); }
template <typename T>
inline T __shfl(T, int, int = 64) { gpumpi_host_unsupported("shfl"); }
template <typename T>
inline T __shfl_xor(T, int, int = 64) { gpumpi_host_unsupported("shfl_xor"); }
template <typename T>
inline T __shfl_up(T, unsigned, int = 64) { gpumpi_host_unsupported("shfl_up"); }
inline void __syncthreads() { gpumpi_host_unsupported("syncthreads"); }

template <typename T>
inline T gpumpi_host_readfirstlane(T) { gpumpi_host_unsupported("readfirstlane"); }
#define __builtin_amdgcn_readfirstlane(v) gpumpi_host_readfirstlane(v)
#define __builtin_amdgcn_readlane(v, l) gpumpi_host_readfirstlane(v)
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
#define __builtin_amdgcn_s_sleep(n) ((void)0)
#define __builtin_amdgcn_endpgm() std::abort()  // never reached by the host checks
#define __builtin_amdgcn_wave_barrier() ((void)0)
// v_alignbyte_b32: bytes [r, r+4) of the 8-byte concatenation lo:hi (hi is the upper half).
inline unsigned gpumpi_host_alignbyte(unsigned hi, unsigned lo, unsigned r) {
    return (unsigned)(((((unsigned long long)hi) << 32) | lo) >> (8 * (r & 3)));
}
#define __builtin_amdgcn_alignbyte(hi, lo, r) gpumpi_host_alignbyte((hi), (lo), (r))
#define __builtin_readsteadycounter() ((uint64_t)0)
#define __builtin_amdgcn_is_private(p) false
#define __builtin_amdgcn_is_shared(p) false

#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __hip_atomic_load(p, o, s) __atomic_load_n((p), (o))
#define __hip_atomic_store(p, v, o, s) __atomic_store_n((p), (v), (o))
#define __hip_atomic_fetch_add(p, v, o, s) __atomic_fetch_add((p), (v), (o))
#define __hip_atomic_compare_exchange_strong(p, e, d, so, fo, s) \
    __atomic_compare_exchange_n((p), (e), (d), false, (so), (fo))

inline float __uint_as_float(unsigned u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline unsigned __float_as_uint(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}
template <typename A, typename B>
inline typename std::common_type<A, B>::type min(A a, B b) { return a < b ? a : b; }
template <typename A, typename B>
inline typename std::common_type<A, B>::type max(A a, B b) { return a < b ? b : a; }
