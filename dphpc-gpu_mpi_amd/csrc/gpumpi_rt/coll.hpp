// Collectives as data-parallel phases over the members' own buffers.
//
// All ranks of the world share one address space, so a collective needs no
// message exchange: every member publishes its buffer addresses in the
// communicator's table, one barrier makes them visible, every member moves
// or folds its share of the data directly between the members' buffers
// (with all waves of its workgroup), and a second barrier ends the
// collective.  HBM traffic is the algorithmic minimum: Allreduce reads each
// input once and writes each output once (2*P*N*s bytes).
//
// Replaces the reference's root-serial p2p collectives and native WORLD
// paths (gpu_libs/gpu_mpi/mpi.cu:86-140 Bcast, :148-235 Reduce, :276-282
// Allreduce, :292-322 Gather, :333-408 Alltoall(v), :410-541 Allgather,
// Gatherv, Scatter(v)).  Semantics are the MPI standard's; reductions keep
// the reference's generic rank-order fold.
#pragma once

#include "ops.hpp"
#include "p2p.hpp"

namespace gpumpi {
namespace rt {

// Collective table words of one member.
enum SlotWord { SW_SBUF = 0, SW_RBUF = 1, SW_SCNT = 2, SW_SDSP = 3, SW_RCNT = 4, SW_RDSP = 5, SW_AUX = 6,
                SW_AUX2 = 7 };

enum SegKind { SEG_ALLTOALLV = 0, SEG_ALLTOALL = 1, SEG_ALLGATHER = 2, SEG_ALLGATHERV = 3 };

// ---------------------------------------------------------------------------
// Collective staging: the second half of the rank's staging slice, bump
// allocated per collective.
// ---------------------------------------------------------------------------
struct CollStage {
    uint64_t top = 0;
    __device__ uint8_t* take(uint64_t bytes) {
        const uint64_t need = (bytes + 255) & ~(uint64_t)255;
        if (top + need > staging_half()) {
            GPUMPI_FAIL(MPI_ERR_NO_MEM, "collective staging space exhausted for a private buffer");
            return nullptr;
        }
        uint8_t* p = staging_base() + staging_half() + top;
        top += need;
        return p;
    }
    // Readable-by-peers copy of `p` (p itself when it is already global).
    __device__ const void* in(const void* p, uint64_t bytes) {
        if (p == nullptr || is_global_ptr(p)) return p;
        uint8_t* s = take(bytes);
        if (s) copy_part(s, p, bytes, lane_id(), 64);
        return s;
    }
    // Peer-writable stand-in for output buffer `p`.
    __device__ void* out(void* p, uint64_t bytes) {
        if (p == nullptr || is_global_ptr(p)) return p;
        return take(bytes);
    }
};

// Copy back a staged private output.
__device__ inline void stage_back(void* user, const void* staged, uint64_t bytes) {
    if (user != staged && user && staged) copy_all_lanes(user, staged, bytes);
}

// ---------------------------------------------------------------------------
// Commands run by every wave of a rank
// ---------------------------------------------------------------------------
// Segment k of a multi-segment copy into this member's receive buffer.
// The copy is min(send block, receive block); a send block larger than the
// receive block sets *trunc (MPI_ERR_TRUNCATE).
__device__ inline void seg_info(const Cmd& c, int k, uint8_t** dst, const uint8_t** src, uint64_t* bytes,
                                bool* trunc) {
    const CollSlot* tab = (const CollSlot*)(uintptr_t)c.a[0];
    const int me = c.i1;
    const uint64_t ses = c.a[1], res = c.a[2];
    const CollSlot& mine = tab[me];
    const CollSlot& peer = tab[k];
    uint8_t* rbuf = (uint8_t*)(uintptr_t)mine.w[SW_RBUF];
    const uint8_t* sbuf = (const uint8_t*)(uintptr_t)peer.w[SW_SBUF];
    switch (c.i2) {
    case SEG_ALLTOALLV: {
        const int* scnt = (const int*)(uintptr_t)peer.w[SW_SCNT];
        const int* sdsp = (const int*)(uintptr_t)peer.w[SW_SDSP];
        const int* rcnt = (const int*)(uintptr_t)mine.w[SW_RCNT];
        const int* rdsp = (const int*)(uintptr_t)mine.w[SW_RDSP];
        uint64_t sb = (uint64_t)scnt[me] * ses, rb = (uint64_t)rcnt[k] * res;
        *bytes = sb < rb ? sb : rb;
        *trunc = sb > rb;
        *src = sbuf + (int64_t)sdsp[me] * (int64_t)ses;
        *dst = rbuf + (int64_t)rdsp[k] * (int64_t)res;
        break;
    }
    case SEG_ALLTOALL: {
        const uint64_t sblk = peer.w[SW_SCNT] * ses, rblk = mine.w[SW_RCNT] * res;
        *bytes = sblk < rblk ? sblk : rblk;
        *trunc = sblk > rblk;
        *src = sbuf + (uint64_t)me * sblk;
        *dst = rbuf + (uint64_t)k * rblk;
        break;
    }
    case SEG_ALLGATHER: {
        const uint64_t sblk = peer.w[SW_SCNT] * ses, rblk = mine.w[SW_RCNT] * res;
        *bytes = sblk < rblk ? sblk : rblk;
        *trunc = sblk > rblk;
        *src = sbuf;
        *dst = rbuf + (uint64_t)k * rblk;
        break;
    }
    default: {  // SEG_ALLGATHERV
        const int* rcnt = (const int*)(uintptr_t)mine.w[SW_RCNT];
        const int* rdsp = (const int*)(uintptr_t)mine.w[SW_RDSP];
        const uint64_t sb = peer.w[SW_SCNT] * ses, rb = (uint64_t)rcnt[k] * res;
        *bytes = sb < rb ? sb : rb;
        *trunc = sb > rb;
        *src = sbuf;
        *dst = rbuf + (int64_t)rdsp[k] * (int64_t)res;
        break;
    }
    }
}

// The segments of a multi-segment copy are visited in the order
// j = 1..P (peer k = (me + j) mod P, spreading the sources).  A wave first
// loads the descriptors of up to 64 of its segments at once, one lane per
// segment (one round trip of table, count and displacement loads instead
// of one per segment), then copies them one after another.
//   large segments: every worker of the rank on each segment in turn;
//   small segments: wave w takes segments j = 1 + w, 1 + w + nw, ...
__device__ inline void run_segments(const Cmd& c, int t, int nt) {
    const int P = c.i0, me = c.i1;
    const bool skip_self = c.i3 != 0;
    const bool large = c.i4 != 0;
    const int lane = t & 63;
    const int first = large ? 1 : 1 + (t >> 6);
    const int step = large ? 1 : (nt + 63) >> 6;
    const int count = first <= P ? (P - first) / step + 1 : 0;
    for (int b = 0; b < count; b += 64) {
        // Lane l: descriptor of this wave's segment b + l.
        uint64_t d = 0, sp = 0, n = 0;
        bool trunc = false;
        const int i = b + lane;
        if (i < count) {
            const int k = (me + first + i * step) % P;
            if (!(k == me && skip_self)) {
                uint8_t* dd;
                const uint8_t* ss;
                seg_info(c, k, &dd, &ss, &n, &trunc);
                d = (uint64_t)(uintptr_t)dd;
                sp = (uint64_t)(uintptr_t)ss;
            }
        }
        // Every worker visits each descriptor once in the large mode, so only
        // the wave that owns worker 0 reports; in the small mode each wave
        // reports its own segments.
        if (ballot(trunc) && (!large || t < 64))
            GPUMPI_FAIL(MPI_ERR_TRUNCATE, "all-to-all/all-gather block larger than the receive block");
        const int nb = count - b < 64 ? count - b : 64;
        for (int s2 = 0; s2 < nb; s2++) {
            const uint64_t bytes = lane_u64(n, s2);
            if (bytes == 0) continue;
            uint8_t* dst = (uint8_t*)(uintptr_t)lane_u64(d, s2);
            const uint8_t* src = (const uint8_t*)(uintptr_t)lane_u64(sp, s2);
            if (large) copy_part(dst, src, bytes, t, nt);
            else copy_part(dst, src, bytes, lane, 64);
        }
    }
}

__device__ inline void run_cmd(const Cmd& c, int t, int nt) {
    switch (c.op) {
    case CMD_COPY:
        copy_part((void*)(uintptr_t)c.a[0], (const void*)(uintptr_t)c.a[1], c.a[2], t, nt);
        break;
    case CMD_REDUCE: {
        const uint64_t* mem = (const uint64_t*)(uintptr_t)c.a[0];  // members_tab()
        FoldArgs A;
        A.ins_w = &mem[0];
        A.outs_w = &mem[1];
        A.stride_words = 2;
        A.nmembers = c.i2;
        A.root = c.i3;
        A.vec = c.i4;
        A.lo = c.a[1];
        A.hi = c.a[2];
        fold_builtin(c.i0, c.i1, A, t, nt);
        break;
    }
    case CMD_ALLTOALLV:
        run_segments(c, t, nt);
        break;
    case CMD_USER: {
        auto fn = (void (*)(const void*, int, int))(uintptr_t)c.a[0];
        fn(&c.a[1], t, nt);
        break;
    }
    default:
        break;
    }
}

// Run a command on wave 0 alone when it is small, on every wave otherwise.
__device__ inline void run_sized(const Cmd& c, uint64_t bytes_moved) {
    if (bytes_moved >= (uint64_t)ctx()->helper_min_bytes && ctx()->waves > 1) {
        wg_run(c);
    } else {
        run_cmd(c, lane_id(), 64);
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Collectives
// ---------------------------------------------------------------------------
struct CollCtx {
    CommShared* c;
    int size;
    int me;
    CollSlot* tab;
};

__device__ inline int coll_begin(MPI_Comm comm, CollCtx* cc) {
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "collective on MPI_COMM_NULL");
    cc->c = comm_get(comm);
    cc->size = cc->c->size;
    cc->me = comm_rank(cc->c);
    cc->tab = cc->c->slots;
    if (cc->me < 0) return GPUMPI_FAIL(MPI_ERR_COMM, "calling rank is not a member of the communicator");
    return MPI_SUCCESS;
}

__device__ inline void publish(CollCtx& cc, int word, uint64_t v) {
    cc.tab[cc.me].w[word] = v;
}
__device__ inline uint64_t peer_word(CollCtx& cc, int r, int word) { return uniform_u64(cc.tab[r].w[word]); }

// Copy every member's `word_in` and `word_out` table entries into
// members_tab() (wave 0; the helpers see it through the command handoff).
__device__ inline void load_members(CollCtx& cc, int word_in, int word_out) {
    uint64_t* m = members_tab();
    for (int r = lane_id(); r < cc.size; r += 64) {
        m[2 * r] = cc.tab[r].w[word_in];
        m[2 * r + 1] = cc.tab[r].w[word_out];
    }
    wave_sync();
}

// Are the given words of every member 16-byte aligned?
__device__ inline bool all_aligned16(CollCtx& cc, int word_a, int word_b, int only_root) {
    bool bad = false;
    for (int r = lane_id(); r < cc.size; r += 64) {
        const uint64_t a = cc.tab[r].w[word_a];
        const uint64_t b = (only_root < 0 || r == only_root) ? cc.tab[r].w[word_b] : 0;
        if ((a | b) & 15) bad = true;
    }
    return ballot(bad) == 0;
}

// Elements [lo, hi) of a reduction with a user-defined operator, in the
// reference's order (mpi.cu:209-225): acc = x_0, then fn(x_r, acc) for
// r = 1 .. P-1, each call as fn(invec, inoutvec, len, datatype).  The fold
// runs in a rank-heap buffer (an input may be an output in place) with the
// slice split across the wave's lanes, each calling fn on its own run of
// elements; the result is then copied to the root's output, or to every
// member's.  Members' buffers are the table's published words.
__device__ inline int user_fold_slice(CollCtx& cc, MPI_User_function* fn, MPI_Datatype dt, int es, uint64_t lo,
                                      uint64_t hi, bool all, int root) {
    const uint64_t n = hi - lo, bytes = n * (uint64_t)es;
    uint8_t* acc = (uint8_t*)rank_malloc(bytes);
    if (!acc) return MPI_ERR_NO_MEM;
    rank_copy(acc, (const uint8_t*)(uintptr_t)peer_word(cc, 0, SW_SBUF) + lo * es, bytes);
    const uint64_t per = (n + 63) / 64;
    const uint64_t e0 = (uint64_t)lane_id() * per < n ? (uint64_t)lane_id() * per : n;
    const uint64_t e1 = e0 + per < n ? e0 + per : n;
    for (int r = 1; r < cc.size; r++) {
        const uint8_t* x = (const uint8_t*)(uintptr_t)peer_word(cc, r, SW_SBUF) + lo * es;
        int len = (int)(e1 - e0);
        MPI_Datatype d = dt;
        if (len > 0) fn((void*)(x + e0 * es), acc + e0 * es, &len, &d);
        wave_sync();
    }
    for (int m = all ? 0 : root; m < (all ? cc.size : root + 1); m++)
        rank_copy((uint8_t*)(uintptr_t)peer_word(cc, m, SW_RBUF) + lo * es, acc, bytes);
    rank_free(acc, bytes);
    return MPI_SUCCESS;
}

__device__ inline int reduce_common(const void* sendbuf, void* recvbuf, int count, MPI_Datatype dt, MPI_Op op,
                                    int root, MPI_Comm comm) {
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    // A derived (contiguous) datatype reduces with a user-defined operator
    // only: the predefined ones are defined on basic datatypes (op_allowed).
    const int es = type_size(dt);
    if (es <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype in reduction");
    if (count < 0) return GPUMPI_FAIL(MPI_ERR_COUNT, "negative count in reduction");
    const uintptr_t opv = (uintptr_t)op;
    const bool builtin = opv >= 1 && opv <= GPUMPI_NUM_BUILTIN_OPS;
    if (op == MPI_OP_NULL) return GPUMPI_FAIL(MPI_ERR_OP, "MPI_OP_NULL in a reduction");
    MPI_User_function* ufn = builtin ? nullptr : ((const ::MPI_Op_impl*)op)->fn;
    if (!builtin && !ufn) return GPUMPI_FAIL(MPI_ERR_OP, "user-defined operator without a function");
    const int opi = builtin ? (int)opv - 1 : -1;
    if (builtin && !op_allowed(dt, opi)) return GPUMPI_FAIL(MPI_ERR_OP, "combination of datatype and operator not allowed");
    const uint64_t bytes = (uint64_t)count * (uint64_t)es;
    const bool all = root < 0;
    const bool i_out = all || cc.me == root;
    const void* src = (sendbuf == MPI_IN_PLACE) ? recvbuf : sendbuf;
    if (cc.size == 1) {
        if (src != recvbuf && (all || cc.me == root)) rank_copy(recvbuf, src, bytes);
        return MPI_SUCCESS;
    }
    CollStage stg;
    const void* s_eff = stg.in(src, bytes);
    void* r_eff = i_out ? stg.out(recvbuf, bytes) : nullptr;
    if (i_out && r_eff != recvbuf && src == recvbuf) r_eff = (void*)s_eff;  // in place, staged
    publish(cc, SW_SBUF, (uint64_t)(uintptr_t)s_eff);
    publish(cc, SW_RBUF, (uint64_t)(uintptr_t)r_eff);
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    // My slice: groups of 16 bytes, split evenly over the members.
    const int V = es >= 16 ? 1 : 16 / es;
    const uint64_t groups = ((uint64_t)count + V - 1) / V;
    const uint64_t g0 = groups * (uint64_t)cc.me / (uint64_t)cc.size;
    const uint64_t g1 = groups * (uint64_t)(cc.me + 1) / (uint64_t)cc.size;
    uint64_t lo = g0 * V, hi = g1 * V;
    if (hi > (uint64_t)count) hi = (uint64_t)count;
    if (lo < hi && ufn) {
        err = user_fold_slice(cc, ufn, dt, es, lo, hi, all, root);
        if (err) return err;
    } else if (lo < hi) {
        Cmd c{};
        c.op = CMD_REDUCE;
        c.i0 = dt;
        c.i1 = opi;
        c.i2 = cc.size;
        c.i3 = all ? -1 : root;
        c.i4 = all_aligned16(cc, SW_SBUF, SW_RBUF, all ? -1 : root) ? 1 : 0;
        load_members(cc, SW_SBUF, SW_RBUF);
        c.a[0] = (uint64_t)(uintptr_t)members_tab();
        c.a[1] = lo;
        c.a[2] = hi;
        const uint64_t moved = (hi - lo) * (uint64_t)es * (uint64_t)(all ? 2 * cc.size : cc.size + 1);
        run_sized(c, moved);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    if (i_out) stage_back(recvbuf, r_eff, bytes);
    return MPI_SUCCESS;
}

__device__ inline int coll_bcast(void* buf, int count, MPI_Datatype dt, int root, MPI_Comm comm) {
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    const int es = type_size(dt);
    if (es <= 0 || count < 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype or count in MPI_Bcast");
    if (root < 0 || root >= cc.size) return GPUMPI_FAIL(MPI_ERR_RANK, "invalid root in MPI_Bcast");
    const uint64_t bytes = (uint64_t)count * (uint64_t)es;
    if (cc.size == 1) return MPI_SUCCESS;
    CollStage stg;
    if (cc.me == root) publish(cc, SW_SBUF, (uint64_t)(uintptr_t)stg.in(buf, bytes));
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    if (cc.me != root) rank_copy(buf, (const void*)(uintptr_t)peer_word(cc, root, SW_SBUF), bytes);
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    return MPI_SUCCESS;
}

// Gather / Gatherv: every member writes its block into the root's buffer.
__device__ inline int coll_gather(bool v, const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                                  void* recvbuf, int recvcount, const int* recvcounts, const int* displs,
                                  MPI_Datatype recvtype, int root, MPI_Comm comm) {
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    if (root < 0 || root >= cc.size) return GPUMPI_FAIL(MPI_ERR_RANK, "invalid root in gather");
    const bool at_root = cc.me == root;
    if (!at_root && sendbuf == MPI_IN_PLACE)
        return GPUMPI_FAIL(MPI_ERR_BUFFER, "MPI_IN_PLACE is valid only at the root of a gather");
    const int ses = sendbuf == MPI_IN_PLACE ? 1 : type_size(sendtype);
    const int res = at_root ? type_size(recvtype) : 1;
    if (ses <= 0 || res <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype in gather");
    CollStage stg;
    uint64_t rtotal = 0;
    void* r_eff = nullptr;
    if (at_root) {
        if (v) {
            for (int r = lane_id(); r < cc.size; r += 64) {
                const uint64_t end = ((uint64_t)displs[r] + (uint64_t)recvcounts[r]) * (uint64_t)res;
                if (end > rtotal) rtotal = end;
            }
            rtotal = wave_max_u64(rtotal);
            publish(cc, SW_RCNT, (uint64_t)(uintptr_t)stg.in(recvcounts, sizeof(int) * (uint64_t)cc.size));
            publish(cc, SW_RDSP, (uint64_t)(uintptr_t)stg.in(displs, sizeof(int) * (uint64_t)cc.size));
        } else {
            rtotal = (uint64_t)cc.size * (uint64_t)recvcount * (uint64_t)res;
            publish(cc, SW_RCNT, (uint64_t)recvcount);
        }
        r_eff = stg.out(recvbuf, rtotal);
        if (r_eff != recvbuf && !is_global_ptr(recvbuf)) copy_part(r_eff, recvbuf, rtotal, lane_id(), 64);
        publish(cc, SW_RBUF, (uint64_t)(uintptr_t)r_eff);
        publish(cc, SW_AUX, (uint64_t)res);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    const bool in_place = at_root && sendbuf == MPI_IN_PLACE;
    if (!in_place) {
        uint8_t* rb = (uint8_t*)(uintptr_t)peer_word(cc, root, SW_RBUF);
        const uint64_t rres = peer_word(cc, root, SW_AUX);
        uint64_t off;
        if (v) {
            const int* rd = (const int*)(uintptr_t)peer_word(cc, root, SW_RDSP);
            off = (uint64_t)(int64_t)uniform_i32(rd[cc.me]) * rres;
        } else {
            off = (uint64_t)cc.me * peer_word(cc, root, SW_RCNT) * rres;
        }
        rank_copy(rb + off, sendbuf, (uint64_t)sendcount * (uint64_t)ses);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    if (at_root) stage_back(recvbuf, r_eff, rtotal);
    return MPI_SUCCESS;
}

// Scatter / Scatterv: every member reads its block from the root's buffer.
__device__ inline int coll_scatter(bool v, const void* sendbuf, int sendcount, const int* sendcounts,
                                   const int* displs, MPI_Datatype sendtype, void* recvbuf, int recvcount,
                                   MPI_Datatype recvtype, int root, MPI_Comm comm) {
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    if (root < 0 || root >= cc.size) return GPUMPI_FAIL(MPI_ERR_RANK, "invalid root in scatter");
    const bool at_root = cc.me == root;
    const int res = type_size(recvtype);
    if (res <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype in scatter");
    CollStage stg;
    if (at_root) {
        const int ses = type_size(sendtype);
        if (ses <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype in scatter");
        uint64_t stotal = 0;
        if (v) {
            for (int r = lane_id(); r < cc.size; r += 64) {
                const uint64_t end = ((uint64_t)displs[r] + (uint64_t)sendcounts[r]) * (uint64_t)ses;
                if (end > stotal) stotal = end;
            }
            stotal = wave_max_u64(stotal);
            publish(cc, SW_SCNT, (uint64_t)(uintptr_t)stg.in(sendcounts, sizeof(int) * (uint64_t)cc.size));
            publish(cc, SW_SDSP, (uint64_t)(uintptr_t)stg.in(displs, sizeof(int) * (uint64_t)cc.size));
        } else {
            stotal = (uint64_t)cc.size * (uint64_t)sendcount * (uint64_t)ses;
            publish(cc, SW_SCNT, (uint64_t)sendcount);
        }
        publish(cc, SW_SBUF, (uint64_t)(uintptr_t)stg.in(sendbuf, stotal));
        publish(cc, SW_AUX, (uint64_t)ses);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    if (!(at_root && recvbuf == MPI_IN_PLACE)) {
        const uint8_t* sb = (const uint8_t*)(uintptr_t)peer_word(cc, root, SW_SBUF);
        const uint64_t ses = peer_word(cc, root, SW_AUX);
        uint64_t off, n;
        if (v) {
            const int* sc = (const int*)(uintptr_t)peer_word(cc, root, SW_SCNT);
            const int* sd = (const int*)(uintptr_t)peer_word(cc, root, SW_SDSP);
            off = (uint64_t)(int64_t)uniform_i32(sd[cc.me]) * ses;
            n = (uint64_t)uniform_i32(sc[cc.me]) * ses;
        } else {
            n = peer_word(cc, root, SW_SCNT) * ses;
            off = (uint64_t)cc.me * n;
        }
        const uint64_t cap = (uint64_t)recvcount * (uint64_t)res;
        if (n > cap) return GPUMPI_FAIL(MPI_ERR_TRUNCATE, "scatter block larger than the receive buffer");
        rank_copy(recvbuf, sb + off, n);
    }
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    return MPI_SUCCESS;
}

// Alltoall(v) and Allgather(v): every member pulls its incoming blocks from
// the other members' send buffers (receiver-driven, one pass over HBM).
__device__ inline int coll_segments(int kind, const void* sendbuf, int sendcount, const int* sendcounts,
                                    const int* sdispls, MPI_Datatype sendtype, void* recvbuf, int recvcount,
                                    const int* recvcounts, const int* rdispls, MPI_Datatype recvtype,
                                    MPI_Comm comm) {
    CollCtx cc;
    int err = coll_begin(comm, &cc);
    if (err) return err;
    const int P = cc.size;
    const bool in_place = sendbuf == MPI_IN_PLACE;
    int ses = in_place ? type_size(recvtype) : type_size(sendtype);
    const int res = type_size(recvtype);
    if (ses <= 0 || res <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype in all-to-all/all-gather");
    if (in_place && (kind == SEG_ALLGATHER || kind == SEG_ALLGATHERV)) sendcount = 0;
    CollStage stg;
    // Receive extent: the lanes read the count and displacement arrays in
    // parallel (one round trip, not one per member).
    uint64_t rtotal = 0;
    if (kind == SEG_ALLTOALLV || kind == SEG_ALLGATHERV) {
        for (int r = lane_id(); r < P; r += 64) {
            const uint64_t end = ((uint64_t)rdispls[r] + (uint64_t)recvcounts[r]) * (uint64_t)res;
            if (end > rtotal) rtotal = end;
        }
        rtotal = wave_max_u64(rtotal);
    } else {
        rtotal = (uint64_t)P * (uint64_t)recvcount * (uint64_t)res;
    }
    void* r_eff = stg.out(recvbuf, rtotal);
    if (r_eff != recvbuf && in_place) copy_part(r_eff, recvbuf, rtotal, lane_id(), 64);
    // Send extent and own contribution.
    const void* s_eff;
    uint64_t sword_cnt = 0;
    // In-place all-to-all: the blocks to send are the receive buffer's own
    // (recvcounts/rdispls, recvtype), so they travel from a copy of it on
    // the rank heap while the peers overwrite the receive buffer; the own
    // block is already in place.
    uint8_t* ip_copy = nullptr;
    if (in_place && (kind == SEG_ALLTOALLV || kind == SEG_ALLTOALL)) {
        ip_copy = (uint8_t*)rank_malloc(rtotal ? rtotal : 1);
        if (!ip_copy) return MPI_ERR_NO_MEM;
        rank_copy(ip_copy, r_eff, rtotal);
        s_eff = ip_copy;
        if (kind == SEG_ALLTOALLV) {
            publish(cc, SW_SCNT, (uint64_t)(uintptr_t)stg.in(recvcounts, sizeof(int) * (uint64_t)P));
            publish(cc, SW_SDSP, (uint64_t)(uintptr_t)stg.in(rdispls, sizeof(int) * (uint64_t)P));
            publish(cc, SW_RCNT, (uint64_t)(uintptr_t)stg.in(recvcounts, sizeof(int) * (uint64_t)P));
            publish(cc, SW_RDSP, (uint64_t)(uintptr_t)stg.in(rdispls, sizeof(int) * (uint64_t)P));
        } else {
            publish(cc, SW_SCNT, (uint64_t)recvcount);
            publish(cc, SW_RCNT, (uint64_t)recvcount);
        }
    } else if (kind == SEG_ALLTOALLV) {
        uint64_t stotal = 0;
        for (int r = lane_id(); r < P; r += 64) {
            const uint64_t end = ((uint64_t)sdispls[r] + (uint64_t)sendcounts[r]) * (uint64_t)ses;
            if (end > stotal) stotal = end;
        }
        stotal = wave_max_u64(stotal);
        s_eff = stg.in(sendbuf, stotal);
        publish(cc, SW_SCNT, (uint64_t)(uintptr_t)stg.in(sendcounts, sizeof(int) * (uint64_t)P));
        publish(cc, SW_SDSP, (uint64_t)(uintptr_t)stg.in(sdispls, sizeof(int) * (uint64_t)P));
        publish(cc, SW_RCNT, (uint64_t)(uintptr_t)stg.in(recvcounts, sizeof(int) * (uint64_t)P));
        publish(cc, SW_RDSP, (uint64_t)(uintptr_t)stg.in(rdispls, sizeof(int) * (uint64_t)P));
    } else if (kind == SEG_ALLTOALL) {
        s_eff = stg.in(sendbuf, (uint64_t)P * (uint64_t)sendcount * (uint64_t)ses);
        sword_cnt = (uint64_t)sendcount;
        publish(cc, SW_SCNT, sword_cnt);
        publish(cc, SW_RCNT, (uint64_t)recvcount);
    } else {
        if (in_place) {
            // My block already sits in my receive buffer.
            const uint64_t off = kind == SEG_ALLGATHER ? (uint64_t)cc.me * recvcount * (uint64_t)res
                                                       : (uint64_t)rdispls[cc.me] * (uint64_t)res;
            s_eff = (const uint8_t*)r_eff + off;
            sword_cnt = kind == SEG_ALLGATHER ? (uint64_t)recvcount : (uint64_t)recvcounts[cc.me];
        } else {
            s_eff = stg.in(sendbuf, (uint64_t)sendcount * (uint64_t)ses);
            sword_cnt = (uint64_t)sendcount;
        }
        publish(cc, SW_SCNT, sword_cnt);
        if (kind == SEG_ALLGATHER) {
            publish(cc, SW_RCNT, (uint64_t)recvcount);
        } else {
            publish(cc, SW_RCNT, (uint64_t)(uintptr_t)stg.in(recvcounts, sizeof(int) * (uint64_t)P));
            publish(cc, SW_RDSP, (uint64_t)(uintptr_t)stg.in(rdispls, sizeof(int) * (uint64_t)P));
        }
    }
    publish(cc, SW_SBUF, (uint64_t)(uintptr_t)s_eff);
    publish(cc, SW_RBUF, (uint64_t)(uintptr_t)r_eff);
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    Cmd c{};
    c.op = CMD_ALLTOALLV;
    c.i0 = P;
    c.i1 = cc.me;
    c.i2 = kind;
    c.i3 = in_place ? 1 : 0;
    c.i4 = (rtotal / (uint64_t)(P > 0 ? P : 1)) >= (256u << 10) ? 1 : 0;
    c.a[0] = (uint64_t)(uintptr_t)cc.tab;
    c.a[1] = (uint64_t)ses;
    c.a[2] = (uint64_t)res;
    run_sized(c, 2 * rtotal);
    if (!comm_barrier(cc.c, cc.size)) return MPI_ERR_OTHER;
    if (ip_copy) rank_free(ip_copy, rtotal ? rtotal : 1);
    stage_back(recvbuf, r_eff, rtotal);
    return MPI_SUCCESS;
}

}  // namespace rt
}  // namespace gpumpi
