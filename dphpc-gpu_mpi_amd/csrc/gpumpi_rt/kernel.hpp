// The persistent rank kernel: one workgroup per MPI rank.
//
// Replaces __gpu_main_caller of gpu_libs/gpu_main/main.cu:102-130 (one
// thread per rank there).  Wave 0 runs the program; waves 1..W-1 serve it.
#pragma once

namespace gpumpi {
namespace rt {

// World rank of workgroup `b`.  With xcd_map the dispatcher's round-robin
// over XCDs (workgroup b runs on XCD b % X) is inverted so ranks
// [k*P/X, (k+1)*P/X) share XCD k and its L2.
__device__ __forceinline__ int rank_of_block(const Config& cfg, int b) {
    const int P = cfg.nranks, X = cfg.num_xcd;
    if (!cfg.xcd_map || X <= 1 || P % X) return b;
    return (b % X) * (P / X) + b / X;
}

__device__ inline void rank_init(World* w) {
    const Config& cfg = w->cfg;
    const LdsLayout lay = lds_layout(cfg.nranks, cfg.max_requests, cfg.user_lds_bytes, cfg.members_in_lds != 0);
    // Zero the whole LDS image with every thread.
    uint32_t* z = (uint32_t*)gpumpi_lds_base;
    for (uint32_t i = threadIdx.x; i < lay.total / 4; i += blockDim.x) z[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        RankLds* L = ctx();
        L->w = w;
        L->slots = w->slots;
        L->chan_head = w->chan_head;
        L->done_words = w->done_words;
        L->comm_world = w->comm_world;
        L->rank = rank_of_block(cfg, (int)blockIdx.x);
        L->members = cfg.members_in_lds ? (uint64_t*)(gpumpi_lds_base + lay.off_members)
                                        : w->members_hbm + 2 * (uint64_t)cfg.nranks * (uint64_t)L->rank;
        L->nranks = cfg.nranks;
        L->waves = cfg.waves;
        L->threads = cfg.threads;
        L->R = cfg.max_requests;
        L->Q = cfg.chan_slots;
        L->slot_bytes = cfg.slot_bytes;
        L->eager_bytes = cfg.eager_bytes;
        L->helper_min_bytes = cfg.helper_min_bytes;
        L->spin_period = cfg.spin_abort_period;
        L->off_reqs = lay.off_reqs;
        L->off_ring = lay.off_ring;
        L->off_rx = lay.off_rx;
        L->off_tx = lay.off_tx;
        L->off_unmatched = lay.off_unmatched;
        L->off_freebits = lay.off_freebits;
        L->off_user = lay.off_user;
        L->lds_bytes = lay.total;
    }
    __syncthreads();
    // Resume the channel cursors of the previous launch (see World::saves).
    const int rank = ctx()->rank;
    const uint8_t* sv = w->saves + (uint64_t)rank * w->save_bytes;
    if (((const uint32_t*)sv)[1] == kSaveMagic) {
        const uint32_t* src = (const uint32_t*)(sv + 64);
        uint32_t* rx = (uint32_t*)rxc();
        const uint32_t nrx = (uint32_t)cfg.nranks * 4, ntx = (uint32_t)cfg.nranks * 2;
        for (uint32_t i = threadIdx.x; i < nrx; i += blockDim.x) rx[i] = src[i];
        uint32_t* tx = (uint32_t*)txc();
        for (uint32_t i = threadIdx.x; i < ntx; i += blockDim.x) tx[i] = src[nrx + i];
        if (threadIdx.x == 0) ctx()->stamp = *(const uint32_t*)sv;
    }
    if (rank == 0 && threadIdx.x < 64) {
        // The communicator pool's allocator state starts empty (wave 0's own
        // stores, ordered before any use by MPI_Init's barrier).
        uint32_t* head = (uint32_t*)w->comm_pool;
        for (uint32_t i = threadIdx.x; i < kPoolHeadBytes / 4; i += 64) head[i] = 0;
    }
    if (threadIdx.x == 0) {
        w->arena_top[rank] = 0;
        if (rank == 0) {  // ordered before any use by MPI_Init's barrier
            __hip_atomic_store(w->comm_pool_top, kPoolHeadBytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(w->next_ctx, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
}

// Wave 0 saves the channel cursors and the stamp at the rank's exit.  A
// message that arrived but was never received would be matched by the next
// launch, so such a rank flags the world for a reset instead.
__device__ inline void rank_save(World* w) {
    bool stale = false;
    for (int src = lane_id(); src < ctx()->nranks; src += 64) {
        const uint32_t h = rxc()[src].head;
        const uint64_t e = ld_relaxed_agent((const uint64_t*)slot_at(my_rank(), src, h));
        if (env_seq(e) == h + 1) stale = true;
    }
    if (ballot(stale) && elected())
        __hip_atomic_store(w->abort_word + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    const Config& cfg = w->cfg;
    uint8_t* sv = w->saves + (uint64_t)my_rank() * w->save_bytes;
    uint32_t* dst = (uint32_t*)(sv + 64);
    const uint32_t* rx = (const uint32_t*)rxc();
    const uint32_t* tx = (const uint32_t*)txc();
    const uint32_t nrx = (uint32_t)cfg.nranks * 4, ntx = (uint32_t)cfg.nranks * 2;
    for (uint32_t i = (uint32_t)lane_id(); i < nrx; i += 64) dst[i] = rx[i];
    for (uint32_t i = (uint32_t)lane_id(); i < ntx; i += 64) dst[nrx + i] = tx[i];
    *(uint32_t*)sv = ctx()->stamp;
    ((uint32_t*)sv)[1] = kSaveMagic;
}

// End the calling rank's program now, as if it returned `code`: the rank's
// exit work (return value, channel cursors, helper waves released) and then
// the program wave itself ends.  Called by the program wave.
[[noreturn]] __device__ inline void rank_exit(int code) {
    World* w = world();
    if (elected()) w->rank_ret[my_rank()] = code;
    wave_sync();
    rank_save(w);
    helpers_exit();
    fence_release_agent();
    __builtin_amdgcn_endpgm();
}

template <int (*Main)(int, char**)>
__device__ inline void rank_body(World* w, int argc, char** argv) {
    rank_init(w);
    if (threadIdx.x < 64) {
        const int ret = Main(argc, argv);
        w->rank_ret[my_rank()] = ret;
        rank_save(w);
        helpers_exit();
    } else {
        helper_loop();
    }
}

}  // namespace rt
}  // namespace gpumpi

// Kernel of a rank program: gpumpi_rank_kernel<prog> is what the host
// launches (one workgroup per rank, dynamic LDS = lds_layout(P, R).total).
template <int (*Main)(int, char**)>
__global__ void __launch_bounds__(1024) gpumpi_rank_kernel(gpumpi::rt::World* w, int argc, char** argv) {
    gpumpi::rt::rank_body<Main>(w, argc, argv);
}

// ---------------------------------------------------------------------------
// Helpers for rank programs
// ---------------------------------------------------------------------------
// Per-rank globals of converted programs: `T var[MAX_RANKS]` indexed by the
// rank (gpu_libs/gpu_main/global_vars.cuh:15 indexed by thread instead).
#ifndef GPU_MPI_MAX_RANKS
#define GPU_MPI_MAX_RANKS 1024
#endif
#define __gpu_global(var) (var[::gpumpi::rt::my_rank()])
__device__ inline int __gpu_max_threads() { return GPU_MPI_MAX_RANKS; }

// Rank heap, wave-uniform (every lane gets the same pointer).
__device__ inline void* gpumpi_malloc(size_t bytes) { return gpumpi::rt::rank_malloc(bytes); }
__device__ inline void gpumpi_free(void* p, size_t bytes) { gpumpi::rt::rank_free(p, bytes); }
// End this rank's program with return value `code` (C's exit in converted
// programs).
[[noreturn]] __device__ inline void gpumpi_rank_exit(int code) { gpumpi::rt::rank_exit(code); }
// The opaque argument block handed to the launch (host C-ABI user_args).
template <typename T>
__device__ inline T* gpumpi_user_args() { return (T*)gpumpi::rt::world()->user_args; }
// This rank's world rank and the wall clock, without MPI calls.
__device__ inline int gpumpi_world_rank() { return gpumpi::rt::my_rank(); }
__device__ inline uint64_t gpumpi_ticks() { return gpumpi::rt::uniform_u64(gpumpi::rt::wall_ticks()); }
__device__ inline bool gpumpi_is_leader_lane() { return gpumpi::rt::elected(); }

// LDS the launch reserved for the rank program (user_lds_bytes of the
// launch configuration); shared by every wave of the rank.
template <typename T>
__device__ inline T* gpumpi_user_lds() {
    return (T*)((uint8_t*)gpumpi::rt::ctx() + gpumpi::rt::ctx()->off_user);
}

// Run `f(t, nt)` on every thread t of the rank's workgroup (nt threads,
// every wave), e.g. the local compute phases of a rank program; returns
// when all threads are done.  `f` is copied by value (at most 72 bytes of
// captures); captured pointers must be global (or user LDS) memory.
template <typename F>
__device__ inline void gpumpi_parallel(const F& f) {
    gpumpi::rt::par_all_waves(f);
}
// Barrier of the rank's own waves is implicit at the end of gpumpi_parallel.
