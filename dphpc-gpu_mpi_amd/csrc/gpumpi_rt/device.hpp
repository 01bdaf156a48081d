// Device-side core of the runtime: wave utilities, scoped atomics, the rank
// workgroup's LDS context, error recording and abort-aware spinning.
//
// A rank is one workgroup.  Wave 0 runs the rank's program with all 64 lanes
// executing the same (wave-uniform) code; waves 1..W-1 are helpers that wait
// in LDS for work posted by wave 0 (helpers.hpp).  Every function in this
// runtime that is called from the program therefore sees wave-uniform
// arguments.  Plain stores of wave-uniform values are made by every lane
// (a store made by one lane only may later be read back by the other lanes
// from their own per-lane view of memory, which the compiler is free to
// forward from registers); read-modify-write atomics and stores other ranks
// poll are made by one elected lane and the result broadcast.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "world.hpp"

namespace gpumpi {
namespace rt {

// ---------------------------------------------------------------------------
// Wave utilities
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ int wave_id() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// Every lane of the wave reaches this point before any goes on.  Free on
// the GPU, whose lanes run in lockstep (it only keeps the compiler from
// moving memory accesses across it); it marks where lanes hand data to each
// other through memory: lane-parallel stores that other lanes of the same
// wave read next.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }
__device__ __forceinline__ uint64_t active_mask() { return __ballot(1); }
// True on exactly one active lane (the lowest).
__device__ __forceinline__ bool elected() {
    return lane_id() == (int)(__builtin_ctzll(active_mask()));
}
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
// Lane `from`'s value of v (`from` wave-uniform; wave-uniform result).
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int from) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, from);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), from);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = uniform_u32((uint32_t)v), hi = uniform_u32((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    return (T*)(uintptr_t)uniform_u64((uint64_t)(uintptr_t)p);
}

// ---------------------------------------------------------------------------
// Scoped atomics and fences (HIP memory model, agent scope = whole GPU)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T ld_relaxed_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_acquire_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_release_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_relaxed_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_relaxed_system(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ T ld_acquire_system(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ void st_release_system(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void fence_acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acq_rel_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }
__device__ __forceinline__ void fence_acquire_wg() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }
__device__ __forceinline__ void fence_release_wg() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }

// ---------------------------------------------------------------------------
// Rank-private state in LDS
// ---------------------------------------------------------------------------
enum ReqKind : uint8_t { REQ_FREE = 0, REQ_SEND = 1, REQ_RECV = 2 };
enum ReqState : uint8_t {
    RS_PENDING = 0,   // send: no slot yet; recv: not matched yet
    RS_WAIT_DONE = 1, // rendezvous send posted, receiver not done copying
    RS_COMPLETE = 2,
    RS_MATCHED = 3,   // recv matched to a rendezvous message; the copy is still to do
};
// RF_DETACHED: the handle was freed by MPI_Request_free; progress releases
// the request once it completes.
enum ReqFlags : uint8_t { RF_SYNC = 1, RF_STAGED = 2, RF_PRIVATE_DST = 4, RF_PROC_NULL = 8, RF_DETACHED = 16 };

struct Req {
    uint8_t kind;
    uint8_t state;
    uint8_t flags;
    uint8_t pad;
    int32_t peer;       // world rank, or -1 for MPI_ANY_SOURCE
    int32_t tag;        // MPI tag or MPI_ANY_TAG
    int32_t ctx;        // communicator context
    uint64_t bytes;     // send length / receive capacity
    uint64_t buf;       // user (or staged) buffer
    uint64_t got;       // receive: bytes delivered
    uint64_t comm;      // MPI_Comm handle (status source translation)
    int32_t st_source;  // status: source (world rank until translated)
    int32_t st_tag;
    int32_t st_error;
    uint32_t stamp;     // posting order stamp; also the rendezvous done value
    int32_t ringpos;    // position in the posting-order ring (validates entries)
    uint32_t rv_val;    // matched rendezvous: value for the sender's done word
    uint64_t rv_src;    // matched rendezvous: sender's buffer
    uint64_t rv_done;   // matched rendezvous: sender's done word
};

struct RxChan {   // receiver view of channel src -> me
    uint32_t head; // oldest sequence not yet consumed
    uint32_t pad;
    uint64_t mask; // consumed slots in [head, head+Q)
};
struct TxChan {   // sender view of channel me -> dst
    uint32_t seq;        // next sequence to send
    uint32_t head_cache; // last receiver head seen
};

// Work a rank's waves perform together (helpers.hpp).
enum CmdOp : int {
    CMD_NONE = 0,
    CMD_EXIT = 1,
    CMD_COPY = 2,
    CMD_REDUCE = 3,        // fold member inputs over a slice
    CMD_ALLTOALLV = 4,
    CMD_GATHER_COPY = 5,
    CMD_USER = 6,
};
struct Cmd {
    int op;
    int i0, i1, i2, i3, i4;
    uint64_t a[10];
};

struct RankLds {
    World* w;
    // World pointers used on every message, kept here (one LDS read instead
    // of an LDS read plus a global load).
    uint8_t* slots;
    uint32_t* chan_head;
    uint32_t* done_words;
    CommShared* comm_world;
    uint64_t* members;  // collective member table: LDS, or this rank's row of World::members_hbm
    int rank;       // world rank
    int nranks;
    int waves;
    int threads;
    int R;          // request pool size
    int Q;          // channel slots
    int slot_bytes;
    int eager_bytes;
    int helper_min_bytes;
    int spin_period;
    int initialized;
    int finalized;
    int next_ctx_local;
    int test_polls;
    // helper protocol
    uint32_t cmd_seq;
    uint32_t ack;
    Cmd cmd;
    // request bookkeeping
    uint32_t stamp;
    int ring_head, ring_tail;
    int unmatched_any;   // unmatched MPI_ANY_SOURCE receives
    int pending_sends;   // sends still waiting for a channel slot
    int pending_copies;  // matched rendezvous receives whose copy is still to do
    uint64_t bounce_top; // staging bump offset for private rendezvous sends
    int bounce_users;
    int detached;        // requests freed by MPI_Request_free that are still active
    uint64_t heap_free;  // first free block of the rank heap: its offset + 1, 0 for none
    // offsets of the variable arrays from the LDS base
    uint32_t off_reqs, off_ring, off_rx, off_tx, off_unmatched, off_freebits, off_user;
    uint32_t lds_bytes;
};

extern __shared__ __attribute__((aligned(16))) uint8_t gpumpi_lds_base[];

__device__ __forceinline__ RankLds* ctx() { return (RankLds*)gpumpi_lds_base; }
__device__ __forceinline__ Req* reqs() { return (Req*)(gpumpi_lds_base + ctx()->off_reqs); }
__device__ __forceinline__ uint16_t* ring() { return (uint16_t*)(gpumpi_lds_base + ctx()->off_ring); }
__device__ __forceinline__ RxChan* rxc() { return (RxChan*)(gpumpi_lds_base + ctx()->off_rx); }
__device__ __forceinline__ TxChan* txc() { return (TxChan*)(gpumpi_lds_base + ctx()->off_tx); }
__device__ __forceinline__ uint16_t* unmatched_src() {
    return (uint16_t*)(gpumpi_lds_base + ctx()->off_unmatched);
}
__device__ __forceinline__ uint64_t* free_bits() {
    return (uint64_t*)(gpumpi_lds_base + ctx()->off_freebits);
}
// Per-collective copy of the members' two buffer addresses (input, output),
// [2 * member + k]: read once from the communicator's table, then by every
// wave of the rank -- from LDS, or, for worlds too large to keep it there
// beside the per-peer channel arrays, from a rank-private HBM row.
__device__ __forceinline__ uint64_t* members_tab() { return ctx()->members; }

// Byte size of the LDS image for a configuration (host and device agree).
__host__ __device__ inline uint32_t align_up(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }
struct LdsLayout {
    uint32_t off_reqs, off_ring, off_rx, off_tx, off_unmatched, off_freebits, off_members, off_user, total;
};
__host__ __device__ inline LdsLayout lds_layout(int nranks, int R, int user_bytes, bool members_in_lds) {
    LdsLayout L;
    uint32_t o = align_up((uint32_t)sizeof(RankLds), 16);
    L.off_reqs = o; o = align_up(o + (uint32_t)(sizeof(Req) * R), 16);
    L.off_ring = o; o = align_up(o + (uint32_t)(2 * R), 16);
    L.off_rx = o; o = align_up(o + (uint32_t)(sizeof(RxChan) * nranks), 16);
    L.off_tx = o; o = align_up(o + (uint32_t)(sizeof(TxChan) * nranks), 16);
    L.off_unmatched = o; o = align_up(o + (uint32_t)(2 * nranks), 16);
    L.off_freebits = o; o = align_up(o + (uint32_t)(8 * ((R + 63) / 64)), 16);
    L.off_members = o; o = align_up(o + (members_in_lds ? (uint32_t)(16 * nranks) : 0u), 16);
    L.off_user = o; o = align_up(o + (uint32_t)user_bytes, 16);
    L.total = o;
    return L;
}

__device__ __forceinline__ World* world() { return ctx()->w; }
__device__ __forceinline__ int my_rank() { return ctx()->rank; }

// ---------------------------------------------------------------------------
// Errors and abort
// ---------------------------------------------------------------------------
__device__ inline bool aborted() {
    int v = 0;
    if (elected()) v = (int)ld_relaxed_agent(world()->dev_abort) | ld_relaxed_system(world()->abort_word);
    return uniform_i32(v) != 0;
}
// The device abort flag alone (a rank's MPI error), without the host word.
__device__ inline bool aborted_on_device() {
    uint32_t v = 0;
    if (elected()) v = ld_relaxed_agent(world()->dev_abort);
    return uniform_u32(v) != 0;
}

__device__ inline void copy_msg(char* dst, const char* msg, int cap) {
    int i = 0;
    for (; i < cap - 1 && msg[i]; i++) dst[i] = msg[i];
    dst[i] = 0;
}

// Records the first error (first writer wins), counts every error and
// raises the abort word so every rank's waits end (errors are fatal, as
// under MPI_ERRORS_ARE_FATAL).  Returns `code`.
__device__ inline __attribute__((noinline)) int raise_error(int code, int line, const char* msg) {
    if (elected()) {
        ErrorRecord* e = world()->err;
        int expected = 0;
        bool first = __hip_atomic_compare_exchange_strong(&e->code, &expected, code, __ATOMIC_RELAXED,
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (first) {
            e->rank = my_rank();
            e->line = line;
            copy_msg(e->msg, msg, (int)sizeof(e->msg));
        }
        __hip_atomic_fetch_add(&e->count, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(world()->dev_abort, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(world()->abort_word, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return code;
}
#define GPUMPI_FAIL(code, msg) ::gpumpi::rt::raise_error((code), __LINE__, (msg))

// Bounded-politeness spin: sleeps between polls and checks the abort word
// every `spin_period` polls.  keep_going() turns false once aborted.
struct Spinner {
    int n = 0;
    __device__ __forceinline__ bool keep_going() {
        if (++n >= ctx()->spin_period) {
            n = 0;
            if (aborted()) return false;
        }
        __builtin_amdgcn_s_sleep(1);
        return true;
    }
    // Same, without the sleep: for a receive's own poll loop, where the
    // poll's memory round trip already paces it and a sleep only adds to
    // the time until an arrived message is seen.
    __device__ __forceinline__ bool keep_polling() {
        if (++n >= ctx()->spin_period) {
            n = 0;
            if (aborted()) return false;
        }
        return true;
    }
};

// Scope guard of the point-to-point entry points: the program's wave leaves
// the call together.  Inside, lanes take different paths (a lane alone
// writes a flag), and on the GPU the wave reconverges by itself; the marker
// costs no instruction there and keeps lane-by-lane execution (the CPU
// simulation of the runtime) in step for the program's next statement.
struct WaveJoin {
    __device__ __forceinline__ ~WaveJoin() { wave_sync(); }
};

// MPI_Wtime clock: s_memrealtime, a constant-rate counter (wall_hz).
__device__ __forceinline__ uint64_t wall_ticks() { return __builtin_readsteadycounter(); }

}  // namespace rt
}  // namespace gpumpi
