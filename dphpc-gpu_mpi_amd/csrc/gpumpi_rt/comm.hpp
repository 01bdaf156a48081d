// Communicators and groups.
//
// An MPI_Comm other than the predefined constants points at a CommShared in
// HBM that every member shares: O(1) rank lookup in both directions
// (replacing the linear group scan of gpu_libs/gpu_mpi/group.cu:110-122),
// its own barrier word and its collective pointer table.  Groups are
// rank-private arrays of world ranks in the rank's device heap.
#pragma once

#include "helpers.hpp"

struct MPI_Comm_impl {};  // opaque: MPI_Comm values point at CommShared
struct MPI_Group_impl {
    int size;
    int ranks[1];  // world ranks, `size` of them
};

namespace gpumpi {
namespace rt {

// ---------------------------------------------------------------------------
// Rank heap (gpumpi_malloc): a bump allocator over the rank's arena slice.
// free() of the most recent block rewinds; other frees are no-ops.
// ---------------------------------------------------------------------------
__device__ inline void* rank_malloc(uint64_t bytes) {
    World* w = world();
    const uint64_t cap = w->cfg.arena_bytes;
    uint64_t* top = &w->arena_top[my_rank()];
    const uint64_t need = (bytes + 255) & ~(uint64_t)255;
    // One lane's view of the top, one lane's store: the block is the same
    // for every lane of the program's wave however its lanes are scheduled.
    const uint64_t off = uniform_u64(*top);
    if (off + need + 16 > cap) {
        GPUMPI_FAIL(MPI_ERR_NO_MEM, "rank heap exhausted");
        return nullptr;
    }
    uint8_t* base = w->arena + (uint64_t)my_rank() * cap;
    if (elected()) *top = off + need;
    wave_sync();
    return base + off;
}

__device__ inline void rank_free(void* p, uint64_t bytes) {
    World* w = world();
    uint8_t* base = w->arena + (uint64_t)my_rank() * w->cfg.arena_bytes;
    uint64_t* top = &w->arena_top[my_rank()];
    const uint64_t need = (bytes + 255) & ~(uint64_t)255;
    const uint64_t t = uniform_u64(*top);
    if ((uint8_t*)p + need == base + t && elected()) *top = (uint64_t)((uint8_t*)p - base);
    wave_sync();
}

// ---------------------------------------------------------------------------
// Handle resolution
// ---------------------------------------------------------------------------
__device__ __forceinline__ CommShared* comm_get(MPI_Comm c) {
    const uintptr_t v = (uintptr_t)c;
    if (v == 1) return ctx()->comm_world;
    if (v == 2) return world()->comm_self + my_rank();
    return (CommShared*)c;
}
__device__ __forceinline__ int comm_size(CommShared* c) { return c->size; }
// This rank's rank in c, or MPI_UNDEFINED.
__device__ __forceinline__ int comm_rank(CommShared* c) {
    if (c->kind == 0) return my_rank();
    if (c->kind == 1) return 0;
    return c->rank_of_world[my_rank()];
}
__device__ __forceinline__ int comm_world_of(CommShared* c, int r) {
    if (c->kind == 0) return r;
    if (c->kind == 1) return my_rank();
    return c->world_of[r];
}
__device__ __forceinline__ int comm_rank_of_world(CommShared* c, int wr) {
    if (c->kind == 0) return wr;
    if (c->kind == 1) return wr == my_rank() ? 0 : -1;
    return c->rank_of_world[wr];
}
__device__ __forceinline__ bool comm_valid(MPI_Comm c) { return c != MPI_COMM_NULL; }

// ---------------------------------------------------------------------------
// Communicator creation.  Members agree through the parent's collective
// table: every member publishes (colour, key); the member that becomes rank
// 0 of a new communicator carves it from the shared pool, takes a fresh
// context id, fills both rank maps and publishes the pointer; everyone picks
// it up after a second parent barrier.
// Follows MPI_Comm_split's ordering, communicator.cu:147-213 (members of a
// colour ordered by key, ties by parent rank) and context-id uniqueness of
// communicator.cu:37-46 (a global counter instead of Allreduce(MAX)).
// ---------------------------------------------------------------------------
__device__ inline uint8_t* comm_pool_alloc(uint64_t bytes) {
    World* w = world();
    bytes = (bytes + 255) & ~(uint64_t)255;
    uint64_t off = 0;
    if (elected()) off = __hip_atomic_fetch_add(w->comm_pool_top, bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    off = uniform_u64(off);
    if (off + bytes > w->cfg.comm_pool_bytes) {
        GPUMPI_FAIL(MPI_ERR_NO_MEM, "communicator pool exhausted");
        return nullptr;
    }
    return w->comm_pool + off;
}

// Split `parent` (this rank has rank `prank` in it) by (color, key).
// Returns the new communicator or MPI_COMM_NULL for MPI_UNDEFINED colour.
// Three parent barriers: (1) colours/keys published, (2) leaders published
// their new communicators, (3) every member entered itself in both maps.
__device__ inline MPI_Comm comm_split_impl(CommShared* parent, int prank, int color, int key) {
    const int psize = parent->size;
    CollSlot* tab = parent->slots;
    tab[prank].w[0] = (uint64_t)(uint32_t)color;
    tab[prank].w[1] = (uint64_t)(uint32_t)key;
    tab[prank].w[2] = 0;
    if (!comm_barrier(parent, psize)) return MPI_COMM_NULL;
    // New rank = members of my colour ordered before me by (key, parent rank);
    // leader = the member of my colour with the smallest (key, parent rank).
    int cnt_less = 0, cnt_size = 0;
    int lkey = 0x7fffffff, lrank = 0x7fffffff;
    for (int q = lane_id(); q < psize; q += 64) {
        const int qc = (int)(uint32_t)tab[q].w[0];
        const int qk = (int)(uint32_t)tab[q].w[1];
        if (qc != color) continue;
        cnt_size++;
        if (qk < key || (qk == key && q < prank)) cnt_less++;
        if (qk < lkey || (qk == lkey && q < lrank)) {
            lkey = qk;
            lrank = q;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cnt_less += __shfl_xor(cnt_less, off);
        cnt_size += __shfl_xor(cnt_size, off);
        const int ok = __shfl_xor(lkey, off), orr = __shfl_xor(lrank, off);
        if (ok < lkey || (ok == lkey && orr < lrank)) {
            lkey = ok;
            lrank = orr;
        }
    }
    const int new_rank = uniform_i32(cnt_less);
    const int new_size = uniform_i32(cnt_size);
    const int leader = uniform_i32(lrank);
    const int P = ctx()->nranks;
    if (color != MPI_UNDEFINED && new_rank == 0) {
        World* w = world();
        const uint64_t bytes = sizeof(CommShared) + sizeof(int) * (uint64_t)new_size +
                               sizeof(int) * (uint64_t)P + sizeof(CollSlot) * (uint64_t)new_size + 512;
        uint8_t* mem = comm_pool_alloc(bytes);
        if (mem) {
            CommShared* nc = (CommShared*)mem;
            int* world_of = (int*)(mem + sizeof(CommShared));
            int* rank_of_world = world_of + new_size;
            CollSlot* slots = (CollSlot*)(((uintptr_t)(rank_of_world + P) + 255) & ~(uintptr_t)255);
            for (int i = lane_id(); i < P; i += 64) rank_of_world[i] = -1;
            int ctxid = 0;
            if (elected()) ctxid = __hip_atomic_fetch_add(w->next_ctx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ctxid = uniform_i32(ctxid);
            if (ctxid > 32767) GPUMPI_FAIL(MPI_ERR_COMM, "communicator context ids exhausted");
            nc->bar = 0;
            nc->ctx = ctxid;
            nc->size = new_size;
            nc->kind = 2;
            nc->world_of = world_of;
            nc->rank_of_world = rank_of_world;
            nc->slots = slots;
            tab[prank].w[2] = (uint64_t)(uintptr_t)nc;
        }
    }
    if (!comm_barrier(parent, psize)) return MPI_COMM_NULL;
    MPI_Comm out = MPI_COMM_NULL;
    if (color != MPI_UNDEFINED) {
        CommShared* nc = (CommShared*)(uintptr_t)uniform_u64(tab[leader].w[2]);
        if (nc) {
            nc->world_of[new_rank] = my_rank();
            nc->rank_of_world[my_rank()] = new_rank;
            out = (MPI_Comm)nc;
        }
    }
    if (!comm_barrier(parent, psize)) return MPI_COMM_NULL;
    return out;
}

}  // namespace rt
}  // namespace gpumpi
