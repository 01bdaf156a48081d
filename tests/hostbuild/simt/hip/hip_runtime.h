// CPU simulation of the HIP device and host runtime, for tests only.
//
// Builds the whole device runtime, the host launcher and the rank programs
// with a host compiler (under AddressSanitizer) and runs a launch on the
// CPU: every GPU lane is a fiber, a wavefront is 64 fibers that meet at
// every wave-wide operation (ballot, shuffles, readfirstlane), a workgroup's
// fibers meet at __syncthreads, and s_sleep yields.  One OS thread runs all
// fibers cooperatively, so memory is sequentially consistent: the
// simulation finds out-of-bounds and wild accesses, wrong indexing and
// protocol deadlocks.  Memory-ordering races between ranks are found by a
// happens-before checker (SIMT_RACE=1, simt.cpp): every scoped atomic and
// fence below reports to it, and every plain load and store of global
// memory is checked against the accesses of other ranks it is not ordered
// with by agent-scope release/acquire synchronisation.
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
#define __shared__
#define GPUMPI_GLOBAL
#define GPUMPI_LDS
#define gpumpi_lds_base (*gpumpi_host_lds)

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace simt {
struct Wave;
struct Block;
struct Fiber {
    void* sp = nullptr;
    char* stack_lo = nullptr;
    size_t stack_size = 0;
    dim3 tid, bid;
    int lane = 0;
    Wave* wave = nullptr;
    Block* block = nullptr;
    bool done = false;
    uint64_t nops = 0;  // wave operations this lane has taken part in
    const uint64_t* wait_gen = nullptr;  // blocked until *wait_gen != wait_val
    uint64_t wait_val = 0;
    bool in_atomic = false;  // inside a scoped atomic: its access is not a plain one
};
extern Fiber* cur;
extern dim3 grid_dim, block_dim;
enum WaveOp { OP_FIRST = 1, OP_SHFL, OP_XOR, OP_UP, OP_BALLOT, OP_SYNC };
uint64_t wave_op(int op, uint64_t v, int arg);
void block_barrier();
void yield();
void mem_step(const void* addr);
void note_alloc(void* p, size_t n);
void note_free(void* p);
bool is_private(const void* p);
void* lane_private(const void* p);
bool is_shared(const void* p);
uint64_t steady_counter();
[[noreturn]] void fail(const char* what);
// The calling lane's fiber ends here, as the lane's part of s_endpgm.
[[noreturn]] void lane_exit();
// Happens-before checker hooks (no-ops unless SIMT_RACE=1).
enum RaceKind { RACE_LOAD = 0, RACE_STORE = 1, RACE_RMW = 2 };
void race_fence(int order, const char* scope);
void race_atomic(const void* p, int kind, int order, int scope);
}  // namespace simt

#define threadIdx (::simt::cur->tid)
#define blockIdx (::simt::cur->bid)
#define blockDim (::simt::block_dim)
#define gridDim (::simt::grid_dim)

struct alignas(16) uint4 {
    unsigned x, y, z, w;
};
struct alignas(8) uint2 {
    unsigned x, y;
};
struct alignas(16) float4 {
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }

template <typename T>
inline uint64_t simt_bits(T v) {
    static_assert(sizeof(T) <= 8, "wave operations move at most 8 bytes");
    uint64_t u = 0;
    std::memcpy(&u, &v, sizeof(T));
    return u;
}
template <typename T>
inline T simt_from(uint64_t u) {
    T v;
    std::memcpy(&v, &u, sizeof(T));
    return v;
}
inline unsigned __lane_id() { return (unsigned)::simt::cur->lane; }
inline unsigned long long __ballot(int p) { return ::simt::wave_op(::simt::OP_BALLOT, p != 0, 0); }
template <typename T>
inline T __shfl(T v, int src, int = 64) { return simt_from<T>(::simt::wave_op(::simt::OP_SHFL, simt_bits(v), src)); }
template <typename T>
inline T __shfl_xor(T v, int m, int = 64) { return simt_from<T>(::simt::wave_op(::simt::OP_XOR, simt_bits(v), m)); }
template <typename T>
inline T __shfl_up(T v, unsigned d, int = 64) { return simt_from<T>(::simt::wave_op(::simt::OP_UP, simt_bits(v), (int)d)); }
inline void __syncthreads() { ::simt::block_barrier(); }
template <typename T>
inline T simt_readfirstlane(T v) { return simt_from<T>(::simt::wave_op(::simt::OP_FIRST, simt_bits(v), 0)); }
#define __builtin_amdgcn_readfirstlane(v) simt_readfirstlane(v)
#define __builtin_amdgcn_readlane(v, l) __shfl((v), (l))
#define __builtin_amdgcn_fence(order, scope) ::simt::race_fence((order), (scope))
#define __builtin_amdgcn_s_sleep(n) ::simt::yield()
#define __builtin_amdgcn_endpgm() ::simt::lane_exit()
#define __builtin_amdgcn_wave_barrier() ((void)::simt::wave_op(::simt::OP_SYNC, 0, 0))
inline unsigned simt_alignbyte(unsigned hi, unsigned lo, unsigned r) {
    return (unsigned)(((((unsigned long long)hi) << 32) | lo) >> (8 * (r & 3)));
}
#define __builtin_amdgcn_alignbyte(hi, lo, r) simt_alignbyte((hi), (lo), (r))
#define __builtin_readsteadycounter() ::simt::steady_counter()
#define __builtin_amdgcn_is_private(p) ::simt::is_private((const void*)(p))
#define GPUMPI_LANE_PRIVATE(p) ((decltype(p))::simt::lane_private((const void*)(p)))
#define __builtin_amdgcn_is_shared(p) ::simt::is_shared((const void*)(p))

#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
// Scoped atomics: the access itself (traced like any load or store, which
// keeps the wave's lockstep) is marked atomic for the checker, which then
// applies the operation's synchronisation.  Nothing yields between the
// access and the report, so the two are one step of the schedule.
template <typename T>
inline T simt_atomic_load(const T* p, int o, int s) {
    ::simt::cur->in_atomic = true;
    const T v = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    ::simt::cur->in_atomic = false;
    ::simt::race_atomic(p, ::simt::RACE_LOAD, o, s);
    return v;
}
template <typename T, typename U>
inline void simt_atomic_store(T* p, U v, int o, int s) {
    ::simt::cur->in_atomic = true;
    __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST);
    ::simt::cur->in_atomic = false;
    ::simt::race_atomic(p, ::simt::RACE_STORE, o, s);
}
template <typename T, typename U>
inline T simt_atomic_fetch_add(T* p, U v, int o, int s) {
    const T r = __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST);
    ::simt::race_atomic(p, ::simt::RACE_RMW, o, s);
    return r;
}
template <typename T, typename U>
inline bool simt_atomic_cas(T* p, T* e, U d, int so, int fo, int s) {
    ::simt::cur->in_atomic = true;  // *e is the caller's own memory
    const bool ok = __atomic_compare_exchange_n(p, e, (T)d, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
    ::simt::cur->in_atomic = false;
    ::simt::race_atomic(p, ::simt::RACE_RMW, ok ? so : fo, s);
    return ok;
}
#define __hip_atomic_load(p, o, s) simt_atomic_load((p), (o), (s))
#define __hip_atomic_store(p, v, o, s) simt_atomic_store((p), (v), (o), (s))
#define __hip_atomic_fetch_add(p, v, o, s) simt_atomic_fetch_add((p), (v), (o), (s))
#define __hip_atomic_compare_exchange_strong(p, e, d, so, fo, s) simt_atomic_cas((p), (e), (d), (so), (fo), (s))
template <typename T, typename U>
inline T atomicAdd(T* p, U v) { return simt_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T, typename U>
inline T atomicSub(T* p, U v) {
    const T r = __atomic_fetch_sub(p, (T)v, __ATOMIC_SEQ_CST);
    ::simt::race_atomic(p, ::simt::RACE_RMW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return r;
}

inline float __uint_as_float(unsigned u) { return simt_from<float>(u); }
inline unsigned __float_as_uint(float f) { return (unsigned)simt_bits(f); }
template <typename A, typename B>
inline typename std::common_type<A, B>::type min(A a, B b) { return a < b ? a : b; }
template <typename A, typename B>
inline typename std::common_type<A, B>::type max(A a, B b) { return a < b ? b : a; }

// ---------------------------------------------------------------------------
// Host API subset used by the launcher and the programs' helper kernels.
// ---------------------------------------------------------------------------
typedef int hipError_t;
enum {
    hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600,
    hipErrorCooperativeLaunchTooLarge = 720, hipErrorNotSupported = 801
};
typedef struct simt_stream* hipStream_t;
typedef struct simt_event* hipEvent_t;
typedef struct simt_graph* hipGraph_t;
typedef struct simt_graph* hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipDeviceAttribute_t {
    hipDeviceAttributeMultiprocessorCount = 1, hipDeviceAttributeWallClockRate, hipDeviceAttributeMaxSharedMemoryPerBlock,
    hipDeviceAttributeMaxSharedMemoryPerMultiprocessor
};
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 1 };
enum hipLimit_t { hipLimitStackSize = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 1 };
#define hipStreamNonBlocking 1u
#define hipHostMallocCoherent 1u

hipError_t hipMalloc(void** p, size_t n);
template <typename T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
template <typename T>
inline hipError_t hipHostMalloc(T** p, size_t n, unsigned flags) { return hipHostMalloc((void**)p, n, flags); }
hipError_t hipHostFree(void* p);
hipError_t hipMemset(void* p, int v, size_t n);
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t s);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDevice(int* d);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int dev);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void* k, int threads, size_t lds);
hipError_t hipFuncSetAttribute(const void* k, hipFuncAttribute a, int v);
hipError_t hipDeviceSetLimit(hipLimit_t l, size_t v);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipLaunchCooperativeKernel(const void* k, dim3 grid, dim3 block, void** args, unsigned lds, hipStream_t s);
hipError_t hipLaunchKernel(const void* k, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t s);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
template <typename T>
inline hipError_t hipFuncSetAttribute(T* k, hipFuncAttribute a, int v) { return hipFuncSetAttribute((const void*)k, a, v); }

namespace simt {
// Plain (non-cooperative) kernels of the programs (fills, checks): run one
// thread after another on the calling thread; they use no wave operations.
void plain_begin(dim3 grid, dim3 block);
void plain_thread(unsigned bx, unsigned by, unsigned t);
void plain_end();
}  // namespace simt
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                      \
    do {                                                                                 \
        const dim3 g_ = dim3(grid), b_ = dim3(block);                                    \
        ::simt::plain_begin(g_, b_);                                                     \
        for (unsigned by_ = 0; by_ < g_.y; by_++)                                        \
            for (unsigned bx_ = 0; bx_ < g_.x; bx_++)                                    \
                for (unsigned t_ = 0; t_ < b_.x; t_++) {                                 \
                    ::simt::plain_thread(bx_, by_, t_);                          \
                    kernel(__VA_ARGS__);                                                 \
                }                                                                        \
        ::simt::plain_end();                                                             \
    } while (0)
