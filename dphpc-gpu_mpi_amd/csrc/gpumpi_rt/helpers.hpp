// Helper-wave protocol.  Wave 0 (the rank's program) posts a Cmd in LDS and
// bumps cmd_seq; waves 1..W-1 wake, run their share, and bump ack.  Only
// LDS flags are used -- no s_barrier -- so the protocol is valid whatever
// control flow of the user program wave 0 is in.
#pragma once

#include "copy.hpp"

namespace gpumpi {
namespace rt {

__device__ void run_cmd(const Cmd& c, int t, int nt);  // coll.hpp

// Helper-wave handoff: post `c` for waves 1..W-1 (returns the sequence to
// wait for), and wait until they have all acknowledged it.
__device__ inline uint32_t wg_post(const Cmd& c) {
    RankLds* L = ctx();
    const uint32_t seq = L->cmd_seq + 1;
    L->cmd = c;
    fence_release_wg();
    __hip_atomic_store(&L->cmd_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    return seq;
}
// Every command is a bounded data-parallel loop that waits on no other
// rank, so this wait always ends and never gives up, not even after an
// abort: wave 0 must not reuse the command, the LDS member table or the
// buffers a command names while a helper may still be reading them.
__device__ inline void wg_wait(uint32_t seq) {
    RankLds* L = ctx();
    const uint32_t target = seq * (uint32_t)(L->waves - 1);
    while (__hip_atomic_load(&L->ack, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < target)
        __builtin_amdgcn_s_sleep(1);
}

// Run `c` on every wave of this rank (wave 0 included) and return when all
// waves are done with it.
__device__ inline void wg_run(const Cmd& c) {
    RankLds* L = ctx();
    const int waves = L->waves;
    if (waves == 1) {
        run_cmd(c, lane_id(), 64);
        wave_sync();
        return;
    }
    const uint32_t seq = wg_post(c);
    run_cmd(c, (int)threadIdx.x, waves * 64);
    wg_wait(seq);
    wave_sync();
}

// A device functor run by every thread of the rank (CMD_USER): the functor
// travels by value inside the LDS command (at most 72 bytes).  It is copied
// into a local before the call: its captures then live in registers, where
// the body's own stores cannot alias them (through the storage pointer
// every store may, and each loop iteration would reload and wait for them).
template <typename F>
__device__ void user_trampoline(const void* storage, int t, int nt) {
    const F f = *(const F*)storage;
    f(t, nt);
}
template <typename F>
__device__ inline void par_all_waves(const F& f) {
    static_assert(sizeof(F) <= 9 * sizeof(uint64_t), "parallel body captures exceed 72 bytes");
    Cmd c{};
    c.op = CMD_USER;
    c.a[0] = (uint64_t)(uintptr_t)&user_trampoline<F>;
    __builtin_memcpy(&c.a[1], &f, sizeof(F));
    wg_run(c);
}
// Same, but on wave 0 alone when the work touches fewer than
// helper_min_bytes of memory.
template <typename F>
__device__ inline void par_run(const F& f, uint64_t bytes) {
    if (bytes < (uint64_t)ctx()->helper_min_bytes || ctx()->waves == 1) {
        f(lane_id(), 64);
        wave_sync();
    } else {
        par_all_waves(f);
    }
}

// Copy by this rank: wave 0 alone for small or private-destination copies,
// every wave for large ones.
__device__ inline void rank_copy(void* dst, const void* src, uint64_t n) {
    if (n == 0) return;
    if (is_private_ptr(dst)) {
        copy_all_lanes(GPUMPI_LANE_PRIVATE(dst), src, n);
        return;
    }
    if (is_private_ptr(src) || n < (uint64_t)ctx()->helper_min_bytes || ctx()->waves == 1) {
        copy_part(dst, src, n, lane_id(), 64);
        return;
    }
    // Wave 0 runs its share of the copy inline (not through run_cmd), so the
    // point-to-point path that calls this stays free of the large command
    // dispatcher and its register demands.
    Cmd c{};
    c.op = CMD_COPY;
    c.a[0] = (uint64_t)(uintptr_t)dst;
    c.a[1] = (uint64_t)(uintptr_t)src;
    c.a[2] = n;
    const uint32_t seq = wg_post(c);
    copy_part(dst, src, n, (int)threadIdx.x, ctx()->waves * 64);
    wg_wait(seq);
}

// Body of waves 1..W-1.  The helpers leave only on CMD_EXIT, which wave 0
// posts when the rank's program returns (also after an abort: every wait
// of wave 0 ends on the abort word), so a command posted after an abort
// still finds its helpers.
__device__ inline void helper_loop() {
    RankLds* L = ctx();
    uint32_t seen = 0;
    for (;;) {
        uint32_t s;
        while ((s = __hip_atomic_load(&L->cmd_seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) == seen)
            __builtin_amdgcn_s_sleep(2);
        seen = s;
        Cmd c = L->cmd;
        if (c.op == CMD_EXIT) return;
        run_cmd(c, (int)threadIdx.x, L->waves * 64);
        fence_release_wg();
        if (elected()) __hip_atomic_fetch_add(&L->ack, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// Wave 0 tells the helpers to leave (end of the rank's program).
__device__ inline void helpers_exit() {
    RankLds* L = ctx();
    if (L->waves == 1) return;
    const uint32_t seq = L->cmd_seq + 1;
    L->cmd.op = CMD_EXIT;
    fence_release_wg();
    __hip_atomic_store(&L->cmd_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ---------------------------------------------------------------------------
// Communicator barrier: a monotonically counting arrival word per
// communicator; barrier k of a member completes when the word reaches
// (k+1)*size.  Replaces the CUDA-internal grid sync of
// gpu_libs/gpu_libc/cuda_mpi.cuh:519-521 and works on every communicator
// (the reference's MPI_Barrier trapped on non-WORLD, mpi.cu:324-331).
// ---------------------------------------------------------------------------
// Returns false when the launch was aborted (an MPI error on some rank, or
// the host watchdog): either the wait gave up, or it completed but a rank
// raised an error first -- a rank that errs and goes on to a later barrier
// (MPI_Finalize, say) adds an arrival that completes its peers' barrier of
// an earlier collective.  Either way the members' table entries are not
// all published, so a collective must not read them and returns.
__device__ inline bool comm_barrier(CommShared* c, int size) {
    if (size <= 1) return true;
    uint64_t old = 0;
    fence_release_agent();
    if (elected()) old = __hip_atomic_fetch_add(&c->bar, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    old = uniform_u64(old);
    const uint64_t target = (old / (uint64_t)size + 1) * (uint64_t)size;
    Spinner sp;
    bool ok = true;
    while (uniform_u64(ld_relaxed_agent(&c->bar)) < target) {
        if (!sp.keep_going()) {
            ok = false;
            break;
        }
    }
    fence_acquire_agent();
    return ok && !aborted_on_device();
}

}  // namespace rt
}  // namespace gpumpi
