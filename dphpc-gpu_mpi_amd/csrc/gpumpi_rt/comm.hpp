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
// Rank heap (gpumpi_malloc / gpumpi_free) over the rank's arena slice, in
// blocks of multiples of 256 B.  Free blocks form a list in address order
// (each holds its size and the next free block in its first 16 bytes) and
// merge with free neighbours; a free block that ends at the top lowers the
// top instead.  Allocation takes the first free block that fits, splitting
// off the rest, else raises the top.  Only the rank's program wave calls
// these: one lane walks the list and the result is made wave-uniform.
// ---------------------------------------------------------------------------
struct HeapFree {
    uint64_t bytes;
    uint64_t next;  // offset + 1 of the next free block, 0 at the end
};

__device__ __forceinline__ uint64_t heap_need(uint64_t bytes) {
    return bytes ? (bytes + 255) & ~(uint64_t)255 : 256;
}

__device__ inline void* rank_malloc(uint64_t bytes) {
    World* w = world();
    const uint64_t cap = w->cfg.arena_bytes;
    uint64_t* top = &w->arena_top[my_rank()];
    uint8_t* base = w->arena + (uint64_t)my_rank() * cap;
    const uint64_t need = heap_need(bytes);
    uint64_t got = 0;  // offset + 1 of a free block taken, 0 for none
    if (elected()) {
        RankLds* L = ctx();
        uint64_t prev = 0, cur = L->heap_free;
        while (cur) {
            HeapFree* f = (HeapFree*)(base + cur - 1);
            if (f->bytes >= need) {
                uint64_t next = f->next;
                if (f->bytes > need) {  // the upper part stays free
                    HeapFree* g = (HeapFree*)(base + cur - 1 + need);
                    g->bytes = f->bytes - need;
                    g->next = next;
                    next = cur + need;
                }
                if (prev) ((HeapFree*)(base + prev - 1))->next = next;
                else L->heap_free = next;
                got = cur;
                break;
            }
            prev = cur;
            cur = f->next;
        }
    }
    got = uniform_u64(got);
    if (got) {
        wave_sync();
        return base + got - 1;
    }
    // One lane's view of the top, one lane's store: the block is the same
    // for every lane of the program's wave however its lanes are scheduled.
    const uint64_t off = uniform_u64(*top);
    if (off + need + 16 > cap) {
        GPUMPI_FAIL(MPI_ERR_NO_MEM, "rank heap exhausted");
        return nullptr;
    }
    if (elected()) *top = off + need;
    wave_sync();
    return base + off;
}

// Return a block of `bytes` (as passed to rank_malloc) at p.  A block that is
// not inside the rank's allocated heap, or that overlaps a free block (freed
// twice), raises MPI_ERR_ARG and is left alone.
__device__ inline void rank_free(void* p, uint64_t bytes) {
    if (!p) return;
    World* w = world();
    uint8_t* base = w->arena + (uint64_t)my_rank() * w->cfg.arena_bytes;
    uint64_t* topp = &w->arena_top[my_rank()];
    const uint64_t need = heap_need(bytes);
    const uint64_t top = uniform_u64(*topp);
    const uint64_t off = (uint64_t)((uint8_t*)p - base);
    if ((uint8_t*)p < base || (off & 255) || off + need > top) {
        GPUMPI_FAIL(MPI_ERR_ARG, "gpumpi_free: block outside the rank heap");
        return;
    }
    int bad = 0;
    if (elected()) {
        RankLds* L = ctx();
        auto F = [&](uint64_t o1) { return (HeapFree*)(base + o1 - 1); };
        uint64_t pp = 0, prev = 0, cur = L->heap_free;
        while (cur && cur - 1 < off) {
            pp = prev;
            prev = cur;
            cur = F(cur)->next;
        }
        if ((prev && prev - 1 + F(prev)->bytes > off) || (cur && off + need > cur - 1)) {
            bad = 1;
        } else {
            uint64_t start, start_pred;
            if (prev && prev - 1 + F(prev)->bytes == off) {  // extends the free block below
                F(prev)->bytes += need;
                start = prev;
                start_pred = pp;
            } else {
                F(off + 1)->bytes = need;
                F(off + 1)->next = cur;
                if (prev) F(prev)->next = off + 1;
                else L->heap_free = off + 1;
                start = off + 1;
                start_pred = prev;
            }
            HeapFree* s = F(start);
            if (cur && start - 1 + s->bytes == cur - 1) {  // and the one above
                s->bytes += F(cur)->bytes;
                s->next = F(cur)->next;
            }
            if (start - 1 + s->bytes == top) {  // ends at the top: lower it
                if (start_pred) F(start_pred)->next = s->next;
                else L->heap_free = s->next;
                *topp = start - 1;
            }
        }
    }
    bad = uniform_i32(bad);
    wave_sync();
    if (bad) GPUMPI_FAIL(MPI_ERR_ARG, "gpumpi_free: block already free");
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
// The pool's lock guards its free list; only communicator creation and
// MPI_Comm_free take it.  One lane of the calling wave holds it.
__device__ inline bool pool_lock(PoolHead* h) {
    Spinner sp;
    for (;;) {
        int got = 0;
        if (elected()) {
            int expect = 0;
            got = __hip_atomic_compare_exchange_strong(&h->lock, &expect, 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) ? 1 : 0;
        }
        if (uniform_i32(got)) return true;
        if (!sp.keep_going()) return false;
    }
}
__device__ inline void pool_unlock(PoolHead* h) {
    if (elected()) __hip_atomic_store(&h->lock, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    wave_sync();
}

// A block of at least `bytes` (multiples of 256 B): the first freed block
// that fits (the part beyond `bytes` split off and kept free when it is 4 KiB
// or more), else the top of the pool is raised.  *got is the block's size.
__device__ inline uint8_t* comm_pool_alloc(uint64_t bytes, uint64_t* got) {
    World* w = world();
    bytes = (bytes + 255) & ~(uint64_t)255;
    PoolHead* h = (PoolHead*)w->comm_pool;
    if (!pool_lock(h)) return nullptr;
    uint64_t off = 0, size = bytes;
    if (elected()) {
        uint64_t prev = 0, cur = h->free_head;
        while (cur) {
            PoolFree* f = (PoolFree*)(w->comm_pool + cur);
            if (f->bytes >= bytes) {
                uint64_t next = f->next;
                size = f->bytes;
                if (size - bytes >= 4096) {
                    PoolFree* g = (PoolFree*)(w->comm_pool + cur + bytes);
                    g->bytes = size - bytes;
                    g->next = next;
                    next = cur + bytes;
                    size = bytes;
                }
                if (prev) ((PoolFree*)(w->comm_pool + prev))->next = next;
                else h->free_head = next;
                off = cur;
                break;
            }
            prev = cur;
            cur = f->next;
        }
    }
    pool_unlock(h);
    off = uniform_u64(off);
    size = uniform_u64(size);
    if (!off) {
        if (elected()) off = __hip_atomic_fetch_add(w->comm_pool_top, bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        off = uniform_u64(off);
        size = bytes;
        if (off + bytes > w->cfg.comm_pool_bytes) {
            GPUMPI_FAIL(MPI_ERR_NO_MEM, "communicator pool exhausted");
            return nullptr;
        }
    }
    *got = size;
    return w->comm_pool + off;
}

// Put a communicator's block back on the pool's free list.
__device__ inline void comm_pool_free(uint8_t* mem, uint64_t bytes) {
    World* w = world();
    PoolHead* h = (PoolHead*)w->comm_pool;
    if (!pool_lock(h)) return;
    if (elected()) {
        PoolFree* f = (PoolFree*)mem;
        f->bytes = bytes;
        f->next = h->free_head;
        h->free_head = (uint64_t)(mem - w->comm_pool);
    }
    pool_unlock(h);
}

// A fresh context id: the counter's next one while it lasts (3 .. kMaxCtx),
// then one that MPI_Comm_free returned; -1 when none is left.
__device__ inline int ctx_alloc() {
    World* w = world();
    int id = 0;
    if (elected()) id = __hip_atomic_fetch_add(w->next_ctx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    id = uniform_i32(id);
    if (id <= kMaxCtx) return id;
    uint32_t* bits = (uint32_t*)(w->comm_pool + kPoolCtxBitsOff);
    id = -1;
    if (elected()) {
        for (int wi = 0; wi < (kMaxCtx + 1) / 32 && id < 0; wi++) {
            uint32_t v = __hip_atomic_load(&bits[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (v) {
                const uint32_t bit = v & (0u - v);
                const uint32_t old =
                    __hip_atomic_fetch_and(&bits[wi], ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (old & bit) {
                    id = wi * 32 + __builtin_ctz(bit);
                    break;
                }
                v = old & ~bit;
            }
        }
    }
    return uniform_i32(id);
}

__device__ inline void ctx_free(int id) {
    if (id < 3 || id > kMaxCtx) return;
    uint32_t* bits = (uint32_t*)(world()->comm_pool + kPoolCtxBitsOff);
    if (elected()) __hip_atomic_fetch_or(&bits[id >> 5], 1u << (id & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    wave_sync();
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
        const uint64_t bytes = sizeof(CommShared) + sizeof(int) * (uint64_t)new_size +
                               sizeof(int) * (uint64_t)P + sizeof(CollSlot) * (uint64_t)new_size + 512;
        uint64_t got = 0;
        uint8_t* mem = comm_pool_alloc(bytes, &got);
        const int ctxid = mem ? ctx_alloc() : -1;
        if (mem && ctxid < 0) {
            comm_pool_free(mem, got);
            GPUMPI_FAIL(MPI_ERR_COMM, "communicator context ids exhausted");
            mem = nullptr;
        }
        if (mem) {
            CommShared* nc = (CommShared*)mem;
            int* world_of = (int*)(mem + sizeof(CommShared));
            int* rank_of_world = world_of + new_size;
            CollSlot* slots = (CollSlot*)(((uintptr_t)(rank_of_world + P) + 255) & ~(uintptr_t)255);
            for (int i = lane_id(); i < P; i += 64) rank_of_world[i] = -1;
            nc->bar = 0;
            nc->pool_bytes = got;
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
