// Point-to-point engine.
//
// Replaces the reference's lock-based matching engine
// (gpu_libs/gpu_libc/cuda_mpi.cu:65-471: isend/irecv, progressStarted/
// Posted/Completed Send/Recv, test, wait) with lock-free per-(src,dst)
// channels in HBM:
//
//   * a channel is a ring of Q slots written only by the sender; every
//     8-byte word of a slot carries the message's sequence number + 1 in
//     its upper half, so the receiver validates each word by itself and
//     an eager message needs no fence on either side;
//   * messages up to the eager size travel inside the slot (the send
//     completes at once); larger ones carry the sender's buffer address
//     (released by the sender before the slot words, acquired by the
//     receiver before its copy) and the receiver copies straight from it
//     in MPI_Wait/MPI_Test, then releases the sender's completion word
//     (rendezvous);
//   * the receiver matches: it scans the slots of one channel (or of every
//     member's channel for MPI_ANY_SOURCE) with one lane per slot and a
//     wavefront ballot, taking the oldest matching message -- MPI's
//     non-overtaking order per (source, context, tag);
//   * the receiver publishes how far it has consumed; the sender reuses a
//     slot only after that.
//
// Rank-private bookkeeping (requests, channel cursors) lives in LDS.
#pragma once

#include "comm.hpp"

namespace gpumpi {
namespace rt {

__device__ __forceinline__ uint8_t* slot_at(int dst, int src, uint32_t seq) {
    RankLds* L = ctx();
    const uint64_t idx = ((uint64_t)dst * (uint64_t)L->nranks + (uint64_t)src) * (uint64_t)L->Q +
                         (uint64_t)(seq & (uint32_t)(L->Q - 1));
    return L->slots + idx * (uint64_t)L->slot_bytes;
}

// ---------------------------------------------------------------------------
// Request pool (LDS).  Handles are (stamp << 16) | (index + 1) so a stale
// handle is detected.
// ---------------------------------------------------------------------------
__device__ __forceinline__ MPI_Request req_handle(int idx) {
    return (MPI_Request)(uintptr_t)(((uint64_t)reqs()[idx].stamp << 16) | (uint64_t)(idx + 1));
}
__device__ __forceinline__ int req_index(MPI_Request h) {
    const uint64_t v = uniform_u64((uint64_t)(uintptr_t)h);
    const int idx = (int)(v & 0xffff) - 1;
    if (idx < 0 || idx >= ctx()->R) return -1;
    if (reqs()[idx].kind == REQ_FREE || reqs()[idx].stamp != (uint32_t)(v >> 16)) return -1;
    return idx;
}

// Ring entry i (absolute position) is live iff its request is in use and
// was pushed at i; entries of freed (and possibly reused) requests are skipped.
__device__ __forceinline__ bool ring_live(int i, int* idx_out) {
    const int idx = ring()[i % ctx()->R];
    *idx_out = idx;
    const Req* q = &reqs()[idx];
    return q->kind != REQ_FREE && q->ringpos == i;
}

__device__ inline void ring_pop_dead() {
    RankLds* L = ctx();
    int idx;
    while (L->ring_head < L->ring_tail && !ring_live(L->ring_head, &idx)) L->ring_head++;
}

// Squeeze dead entries out of the ring, keeping posting order.
__device__ inline __attribute__((noinline)) void ring_compact() {
    RankLds* L = ctx();
    const int R = L->R;
    int out = L->ring_head;
    for (int i = L->ring_head; i < L->ring_tail; i++) {
        int idx;
        if (ring_live(i, &idx)) {
            ring()[out % R] = (uint16_t)idx;
            reqs()[idx].ringpos = out;
            out++;
        }
    }
    L->ring_tail = out;
}

__device__ inline void progress(bool copies = true);

__device__ inline __attribute__((noinline)) int req_alloc() {
    RankLds* L = ctx();
    const int R = L->R;
    const int nwords = (R + 63) / 64;
    for (int attempt = 0;; attempt++) {
        ring_pop_dead();
        if (L->ring_tail - L->ring_head >= R) ring_compact();
        if (L->ring_tail - L->ring_head < R) {
            for (int wi = 0; wi < nwords; wi++) {
                uint64_t bits = free_bits()[wi];
                if (wi == nwords - 1 && (R & 63)) bits |= ~((1ull << (R & 63)) - 1);
                if (bits != ~0ull) {
                    const int b = __builtin_ctzll(~bits);
                    const int idx = wi * 64 + b;
                    free_bits()[wi] = free_bits()[wi] | (1ull << b);
                    Req* q = &reqs()[idx];
                    q->state = RS_PENDING;
                    q->flags = 0;
                    uint32_t st = L->stamp + 1;
                    if (st == 0) st = 1;
                    L->stamp = st;
                    q->stamp = st;
                    q->ringpos = L->ring_tail;
                    ring()[L->ring_tail % R] = (uint16_t)idx;
                    L->ring_tail++;
                    return idx;
                }
            }
        }
        if (attempt > 64 || aborted()) break;
        progress(false);
    }
    GPUMPI_FAIL(MPI_ERR_REQUEST, "request pool exhausted (raise pending requests, -e)");
    return -1;
}

__device__ __forceinline__ void req_release(int idx) {
    Req* q = &reqs()[idx];
    q->kind = REQ_FREE;
    free_bits()[idx >> 6] = free_bits()[idx >> 6] & ~(1ull << (idx & 63));
}

// ---------------------------------------------------------------------------
// Staging (bounce) buffers for private / LDS buffers.  The first half of the
// rank's staging slice backs rendezvous sends, the second half collectives.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint8_t* staging_base() {
    World* w = world();
    return w->staging + (uint64_t)my_rank() * w->cfg.staging_bytes;
}
__device__ __forceinline__ uint64_t staging_half() { return world()->cfg.staging_bytes / 2; }

__device__ inline __attribute__((noinline)) uint8_t* bounce_alloc(uint64_t bytes) {
    RankLds* L = ctx();
    const uint64_t need = (bytes + 255) & ~(uint64_t)255;
    if (L->bounce_users == 0) L->bounce_top = 0;
    if (L->bounce_top + need > staging_half()) {
        GPUMPI_FAIL(MPI_ERR_NO_MEM, "staging space for a private-buffer send exhausted");
        return nullptr;
    }
    uint8_t* p = staging_base() + L->bounce_top;
    L->bounce_top += need;
    L->bounce_users++;
    return p;
}

// ---------------------------------------------------------------------------
// Send side
// ---------------------------------------------------------------------------
// Payload word k of an n-byte buffer (bytes [4k, 4k+4), zero padded past n).
__device__ __forceinline__ uint32_t slot_payload_get(const uint8_t* p, uint64_t n, uint64_t k) {
    const uint64_t o = 4 * k;
    if (o + 4 <= n && ((uintptr_t)p & 3) == 0) return *(const uint32_t*)(p + o);
    uint32_t v = 0;
    for (int b = 0; b < 4 && o + (uint64_t)b < n; b++) v |= (uint32_t)p[o + b] << (8 * b);
    return v;
}
// Store payload word k into an n-byte buffer (bytes past n are dropped).
__device__ __forceinline__ void slot_payload_put(uint8_t* p, uint64_t n, uint64_t k, uint32_t v) {
    const uint64_t o = 4 * k;
    if (o + 4 <= n && ((uintptr_t)p & 3) == 0) {
        *(uint32_t*)(p + o) = v;
        return;
    }
    for (int b = 0; b < 4 && o + (uint64_t)b < n; b++) p[o + b] = (uint8_t)(v >> (8 * b));
}
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int l) { return (uint32_t)__shfl((int)v, l); }

// Sequence number of the next free slot of channel me->dst, or false when
// all Q slots are still unconsumed.
__device__ __forceinline__ bool slot_reserve(int dst, uint32_t* seq) {
    RankLds* L = ctx();
    TxChan* tc = &txc()[dst];
    const uint32_t s = tc->seq;
    const uint32_t Q = (uint32_t)L->Q;
    if (s - tc->head_cache >= Q) {
        // The receiver publishes its head only after it has read the slots
        // it consumed, so no fence is needed before reusing them.
        const uint32_t h = uniform_u32(ld_relaxed_agent(&L->chan_head[(uint64_t)dst * L->nranks + L->rank]));
        tc->head_cache = h;
        if (s - h >= Q) return false;
    }
    *seq = s;
    return true;
}

// Write an eager message into slot `seq` of channel me->dst.  Every word of
// the slot (envelope, info, payload) is one 8-byte store tagged with seq+1,
// written by the lanes in parallel with no fence: the receiver validates
// each word by its tag.
__device__ __forceinline__ void slot_write_eager(int dst, uint32_t seq, const void* buf, uint64_t n, int ctxid,
                                                 int tag) {
    RankLds* L = ctx();
    uint64_t* slot = (uint64_t*)slot_at(dst, L->rank, seq);
    const uint32_t seq1 = seq + 1;
    const uint8_t* src = (const uint8_t*)buf;
    const int nw = eager_words(n);
    for (int w = lane_id(); w < nw; w += 64) {
        uint32_t v;
        if (w == SF_ENV) v = slot_word_value(pack_env(0, ctxid, tag));
        else if (w == SF_INFO) v = (SLOT_EAGER << 24) | (uint32_t)n;
        else v = slot_payload_get(src, n, (uint64_t)(w - SF_PAYLOAD));
        st_relaxed_agent(&slot[w], slot_word(seq1, v));
    }
    txc()[dst].seq = seq + 1;
}

// Post send request idx into its channel if a slot is free.  A rendezvous
// send first releases its buffer, which the receiver will read directly,
// then announces address, length and its completion word.
__device__ inline bool try_post_send(int idx) {
    RankLds* L = ctx();
    Req* q = &reqs()[idx];
    const int dst = q->peer, me = L->rank;
    uint32_t s;
    if (!slot_reserve(dst, &s)) return false;
    const uint64_t n = q->bytes;
    const bool eager = n <= (uint64_t)L->eager_bytes && !(q->flags & RF_SYNC);
    if (eager) {
        slot_write_eager(dst, s, (const void*)(uintptr_t)q->buf, n, q->ctx, q->tag);
        q->state = RS_COMPLETE;
        return true;
    }
    uint64_t* slot = (uint64_t*)slot_at(dst, me, s);
    const uint32_t seq1 = s + 1;
    const int lane = lane_id();
    fence_release_agent();
    const uint64_t done = (uint64_t)(uintptr_t)&L->done_words[(uint64_t)me * L->R + idx];
    if (lane < SF_RNDV_WORDS) {
        uint32_t v = 0;
        switch (lane) {
        case SF_ENV: v = slot_word_value(pack_env(0, q->ctx, q->tag)); break;
        case SF_INFO: v = (SLOT_RNDV << 24) | (uint32_t)(n & 0xffffffu); break;
        case SF_SRC_LO: v = (uint32_t)q->buf; break;
        case SF_SRC_HI: v = (uint32_t)(q->buf >> 32); break;
        case SF_DONE_LO: v = (uint32_t)done; break;
        case SF_DONE_HI: v = (uint32_t)(done >> 32); break;
        case SF_DONE_VAL: v = q->stamp; break;
        default: v = (uint32_t)(n >> 24); break;
        }
        st_relaxed_agent(&slot[lane], slot_word(seq1, v));
    }
    txc()[dst].seq = s + 1;
    q->state = RS_WAIT_DONE;
    return true;
}

// Start a send of `bytes` at `buf` to world rank `dst`.  Returns the request
// index, or -1 after an error.
__device__ inline __attribute__((noinline)) int start_send(const void* buf, uint64_t bytes, int dst, int tag, int ctxid,
                                 MPI_Comm comm, bool sync) {
    RankLds* L = ctx();
    const int idx = req_alloc();
    if (idx < 0) return -1;
    Req* q = &reqs()[idx];
    q->kind = REQ_SEND;
    q->state = RS_PENDING;
    q->flags = sync ? RF_SYNC : 0;
    q->peer = dst;
    q->tag = tag;
    q->ctx = ctxid;
    q->bytes = bytes;
    q->buf = (uint64_t)(uintptr_t)buf;
    q->comm = (uint64_t)(uintptr_t)comm;
    q->st_error = MPI_SUCCESS;
    if (dst == MPI_PROC_NULL) {
        q->state = RS_COMPLETE;
        return idx;
    }
    const bool eager = bytes <= (uint64_t)L->eager_bytes && !sync;
    if (!eager && !is_global_ptr(buf)) {
        uint8_t* b = bounce_alloc(bytes);
        if (!b) {
            q->state = RS_COMPLETE;
            q->st_error = MPI_ERR_NO_MEM;
            return idx;
        }
        copy_part(b, buf, bytes, lane_id(), 64);
        q->buf = (uint64_t)(uintptr_t)b;
        q->flags |= RF_STAGED;
    }
    // Non-overtaking: a send may not take a slot while an earlier send to
    // the same destination is still waiting for one.
    bool earlier_waiting = false;
    if (L->pending_sends > 0) {
        for (int i = L->ring_head; i < L->ring_tail && !earlier_waiting; i++) {
            int oi;
            if (!ring_live(i, &oi) || oi == idx) continue;
            const Req* o = &reqs()[oi];
            if (o->kind == REQ_SEND && o->state == RS_PENDING && o->peer == dst) earlier_waiting = true;
        }
    }
    if (earlier_waiting || !try_post_send(idx)) L->pending_sends++;
    return idx;
}

// ---------------------------------------------------------------------------
// Receive side
// ---------------------------------------------------------------------------
// Make lane l's word (index base + l, loaded into v when loaded) valid for
// every word below `need`: re-load those that do not carry tag seq1 yet.
// The envelope was seen already and the other words were stored by the
// same sender instructions, so this waits at most for their propagation.
__device__ inline uint64_t slot_validate(const uint64_t* slot, int base, int need, uint32_t seq1, uint64_t v) {
    const int w = base + lane_id();
    for (;;) {
        const bool bad = w < need && slot_word_seq(v) != seq1;
        if (ballot(bad) == 0) return v;
        if (bad) v = ld_relaxed_agent(&slot[w]);
    }
}

// What reading a slot yields: bytes delivered (eager) or announced
// (rendezvous, with the sender's buffer and completion word), and the error.
struct Taken {
    uint64_t got;
    int err;
    bool rndv;
    uint64_t rv_src, rv_done;
    uint32_t rv_val;
};

// Read message `seq` of channel src->me: copy an eager payload into `dstv`
// (capacity `cap` bytes; a larger message is truncated with
// MPI_ERR_TRUNCATE), or return the rendezvous announcement; then mark the
// slot consumed.  `pre`, when has_pre, is lane l's word l of the slot,
// loaded by the matcher in the same round trip as the envelope.
__device__ inline void slot_take(int src, uint32_t seq, void* dstv, uint64_t cap, bool has_pre, uint64_t pre,
                                 Taken* t) {
    RankLds* L = ctx();
    const int me = L->rank;
    const uint64_t* slot = (const uint64_t*)slot_at(me, src, seq);
    const uint32_t seq1 = seq + 1;
    const int swords = L->slot_bytes / 8;
    // Words 0..63 (at most the slot) in one round trip; the info word then
    // says how many of them (and of later chunks) the message uses.
    const int first = swords < 64 ? swords : 64;
    uint64_t v = has_pre ? pre : (lane_id() < first ? ld_relaxed_agent(&slot[lane_id()]) : 0);
    v = slot_validate(slot, 0, SF_INFO + 1, seq1, v);
    const uint32_t info = lane_u32(slot_word_value(v), SF_INFO);
    const bool eager = ((info >> 24) & SLOT_EAGER) != 0;
    const int need = eager ? eager_words(info & 0xffffffu) : (int)SF_RNDV_WORDS;
    v = slot_validate(slot, 0, need < first ? need : first, seq1, v);
    uint64_t bytes = info & 0xffffffu;
    if (!eager) bytes |= (uint64_t)lane_u32(slot_word_value(v), SF_BYTES_HI) << 24;
    t->got = bytes;
    t->err = MPI_SUCCESS;
    if (bytes > cap) {
        t->got = cap;
        t->err = MPI_ERR_TRUNCATE;
    }
    t->rndv = !eager;
    if (eager) {
        uint8_t* dst = (uint8_t*)dstv;
        const uint64_t n = t->got;
        const bool priv = is_private_ptr(dst);
        for (int base = 0; base < need; base += 64) {
            if (base > 0) {
                const int w = base + lane_id();
                v = w < need ? ld_relaxed_agent(&slot[w]) : 0;
                v = slot_validate(slot, base, need, seq1, v);
            }
            const uint32_t val = slot_word_value(v);
            if (priv) {
                // Every lane fills its own copy of a private buffer.
                uint8_t* own = GPUMPI_LANE_PRIVATE(dst);
                const int lim = need - base < 64 ? need - base : 64;
                for (int j = 0; j < lim; j++) {
                    const uint32_t x = lane_u32(val, j);
                    if (base + j >= SF_PAYLOAD) slot_payload_put(own, n, (uint64_t)(base + j - SF_PAYLOAD), x);
                }
            } else {
                const int w = base + lane_id();
                if (w >= SF_PAYLOAD && w < need) slot_payload_put(dst, n, (uint64_t)(w - SF_PAYLOAD), val);
            }
        }
        wave_sync();
    } else {
        const uint32_t val = slot_word_value(v);
        t->rv_src = ((uint64_t)lane_u32(val, SF_SRC_HI) << 32) | lane_u32(val, SF_SRC_LO);
        t->rv_done = ((uint64_t)lane_u32(val, SF_DONE_HI) << 32) | lane_u32(val, SF_DONE_LO);
        t->rv_val = lane_u32(val, SF_DONE_VAL);
    }
    // Mark consumed; publish the new head once the oldest slots are done.
    // Every word of the slot has been read (validated above), so the sender
    // may reuse it as soon as it sees the head: no fence is needed.
    RxChan* rc = &rxc()[src];
    const uint32_t h0 = rc->head;
    uint64_t mask = rc->mask | (1ull << (seq - h0));
    uint32_t h = h0;
    while (mask & 1ull) {
        mask >>= 1;
        h++;
    }
    rc->mask = mask;
    rc->head = h;
    if (h != h0 && elected()) st_relaxed_agent(&L->chan_head[(uint64_t)me * L->nranks + src], h);
}

// Copy a rendezvous message out of the sender's buffer (every wave of the
// rank), then release the sender's completion word.
__device__ inline __attribute__((noinline)) void rendezvous_copy(void* dst, uint64_t src, uint64_t n,
                                                                 uint64_t done, uint32_t done_val) {
    fence_acquire_agent();  // pairs with the sender's release before it posted
    rank_copy(dst, (const void*)(uintptr_t)src, n);
    fence_release_agent();
    if (elected()) st_relaxed_agent((uint32_t*)(uintptr_t)done, done_val);
}

// Deliver message `seq` of channel src->me to receive request `idx`.  A
// rendezvous message is only matched here: the copy runs later, from a wait
// or test call (finish_rendezvous), so that an MPI_Irecv that finds its
// message already announced returns at once and the rank's own sends get
// posted before it spends time copying.
__device__ inline __attribute__((noinline)) void deliver(int idx, int src, uint32_t seq, uint64_t env, bool has_pre = false,
                               uint64_t pre = 0) {
    RankLds* L = ctx();
    Req* q = &reqs()[idx];
    Taken t;
    slot_take(src, seq, (void*)(uintptr_t)q->buf, q->bytes, has_pre, pre, &t);
    if (t.rndv) {
        q->rv_src = t.rv_src;
        q->rv_done = t.rv_done;
        q->rv_val = t.rv_val;
        L->pending_copies++;
    }
    q->state = t.rndv ? RS_MATCHED : RS_COMPLETE;
    q->got = t.got;
    q->st_source = src;
    q->st_tag = env_tag(env);
    q->st_error = t.err;
    if (q->peer < 0) L->unmatched_any--;
    else unmatched_src()[q->peer] = (uint16_t)(unmatched_src()[q->peer] - 1);
    if (t.err != MPI_SUCCESS) GPUMPI_FAIL(t.err, "message truncated: receive buffer too small");
}

// Slots of one channel are read by different lanes with independent loads,
// so a later slot can be seen filled while an earlier one (written first by
// the sender) still reads as empty.  Matching is therefore confined to the
// prefix of slots seen arrived or already consumed: everything past the
// first slot not yet seen waits for the next poll, which keeps the
// non-overtaking order of MPI.  `group` is the ballot of arrived-or-consumed
// slots of this channel (bit j = slot head+j).
__device__ __forceinline__ uint64_t arrived_prefix(uint64_t group, int Q) {
    const uint64_t all = Q >= 64 ? ~0ull : ((1ull << Q) - 1);
    const uint64_t missing = ~group & all;
    return missing ? ((1ull << __builtin_ctzll(missing)) - 1) : all;
}

// A message found for a receive: channel `src`, sequence `seq`, envelope.
struct Found {
    int src;
    uint32_t seq;
    uint64_t env;
    bool has_pre;  // pre holds this lane's word of the matched slot's first chunk
    uint64_t pre;
};

// Look for the oldest message of channel src->me with context ctxid and tag
// `tag` (or any tag).
__device__ inline bool scan_channel(int src, int ctxid, int tag, Found* f) {
    RankLds* L = ctx();
    const int Q = L->Q;
    const RxChan rc = rxc()[src];
    const int j = lane_id();
    // The oldest unconsumed slot (the head) must have arrived before any
    // slot can match (arrived prefix, below), so a poll first reads only the
    // head slot's first 16 words -- one 128-byte line, not the Q envelopes
    // and a 512-byte chunk.  When the head holds the match (the in-order
    // case) that is the whole poll; words past these are loaded by
    // slot_take's validation.  A head for another context or tag falls
    // through to the scan of the whole window.
    {
        const uint64_t* head_slot = (const uint64_t*)slot_at(L->rank, src, rc.head);
        const int hw = L->slot_bytes / 8 < 16 ? L->slot_bytes / 8 : 16;
        const uint64_t hv = j < hw ? ld_relaxed_agent(&head_slot[j]) : 0;
        const uint64_t env0 = uniform_u64(lane_u64(hv, 0));
        if (env_seq(env0) != rc.head + 1u) return false;
        if (env_ctx(env0) == ctxid && (tag == MPI_ANY_TAG || env_tag(env0) == tag)) {
            f->src = src;
            f->seq = rc.head;
            f->env = env0;
            f->has_pre = true;
            f->pre = hv;
            return true;
        }
    }
    bool ok = false, seen = false;
    uint64_t env = 0;
    // Alongside the envelopes, the first chunk of the oldest slot: when that
    // slot is the match (the common, in-order case) delivery needs no
    // further round trip.
    const uint64_t* head_slot = (const uint64_t*)slot_at(L->rank, src, rc.head);
    const int hw = L->slot_bytes / 8 < 64 ? L->slot_bytes / 8 : 64;
    if (j < Q) {
        if ((rc.mask >> j) & 1ull) {
            seen = true;
        } else {
            env = ld_relaxed_agent((const uint64_t*)slot_at(L->rank, src, rc.head + (uint32_t)j));
        }
    }
    const uint64_t hv = j < hw ? ld_relaxed_agent(&head_slot[j]) : 0;
    if (j < Q && !seen) {
        seen = env_seq(env) == rc.head + (uint32_t)j + 1u;
        ok = seen && env_ctx(env) == ctxid && (tag == MPI_ANY_TAG || env_tag(env) == tag);
    }
    const uint64_t b = ballot(ok) & arrived_prefix(ballot(seen), Q);
    if (!b) return false;
    const int first = __builtin_ctzll(b);
    f->src = src;
    f->seq = rc.head + (uint32_t)first;
    f->env = uniform_u64(__shfl(env, first));
    f->has_pre = first == 0;
    f->pre = hv;
    return true;
}

// Look for the oldest message of channel src->me matching request idx.
__device__ __forceinline__ bool match_from(int idx, int src, Found* f) {
    const Req* q = &reqs()[idx];
    return scan_channel(src, q->ctx, q->tag, f);
}

// MPI_ANY_SOURCE: scan the channels of every member of the communicator,
// lowest member first.  Q divides 64, so each channel's Q slots fall in one
// 64-lane batch (lanes [g*Q, g*Q+Q) of it).
__device__ inline bool match_any(int idx, Found* f) {
    RankLds* L = ctx();
    const Req* q = &reqs()[idx];
    CommShared* c = comm_get((MPI_Comm)(uintptr_t)q->comm);
    const int size = c->size;
    const int Q = L->Q;
    const int total = size * Q;
    const int lane = lane_id();
    const int gbase = lane - lane % Q;  // first lane of this lane's channel
    for (int base = 0; base < total; base += 64) {
        const int k = base + lane;
        bool ok = false, seen = false;
        uint64_t env = 0;
        int src = 0;
        uint32_t seq = 0;
        if (k < total) {
            const int r = k / Q, j = k % Q;
            src = comm_world_of(c, r);
            const RxChan rc = rxc()[src];
            seq = rc.head + (uint32_t)j;
            if ((rc.mask >> j) & 1ull) {
                seen = true;
            } else {
                env = ld_relaxed_agent((const uint64_t*)slot_at(L->rank, src, seq));
                seen = env_seq(env) == seq + 1u;
                ok = seen && env_ctx(env) == q->ctx && (q->tag == MPI_ANY_TAG || env_tag(env) == q->tag);
            }
        }
        const uint64_t seen_b = ballot(seen);
        const uint64_t pre = arrived_prefix(Q >= 64 ? seen_b : (seen_b >> gbase), Q);
        ok = ok && ((pre >> (lane - gbase)) & 1ull);
        const uint64_t b = ballot(ok);
        if (b) {
            const int first = __builtin_ctzll(b);
            f->env = uniform_u64(__shfl(env, first));
            f->src = uniform_i32(__shfl(src, first));
            f->seq = uniform_u32(__shfl((int)seq, first));
            f->has_pre = false;
            f->pre = 0;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ bool find_match(int idx, Found* f) {
    const Req* q = &reqs()[idx];
    return q->peer >= 0 ? match_from(idx, q->peer, f) : match_any(idx, f);
}

// Match and deliver request idx (used when no earlier posted receive could
// take the same message).
__device__ __forceinline__ bool try_match(int idx) {
    Found f;
    if (!find_match(idx, &f)) return false;
    deliver(idx, f.src, f.seq, f.env, f.has_pre, f.pre);
    return true;
}

// The earliest pending receive posted before ring position `pos` that would
// also accept message f (same source or MPI_ANY_SOURCE, same context, tag
// or MPI_ANY_TAG), or -1.  A message that arrives while a progress pass is
// between an earlier receive and a later one must still go to the earlier
// one (MPI's matching order of posted receives).
__device__ inline int earlier_taker(int pos, const Found& f) {
    RankLds* L = ctx();
    // The receive that found f is one of the unmatched receives counted here.
    if (L->unmatched_any + (int)unmatched_src()[f.src] < 2) return -1;
    const int ctxid = env_ctx(f.env), tag = env_tag(f.env);
    for (int i = L->ring_head; i < pos; i++) {
        int oi;
        if (!ring_live(i, &oi)) continue;
        const Req* o = &reqs()[oi];
        if (o->kind != REQ_RECV || o->state != RS_PENDING) continue;
        if ((o->peer != f.src && o->peer >= 0) || o->ctx != ctxid) continue;
        if (o->tag != MPI_ANY_TAG && o->tag != tag) continue;
        return oi;
    }
    return -1;
}

// Copy a matched rendezvous message out of the sender's buffer (every wave
// of the rank), then release the sender's done word.
__device__ inline __attribute__((noinline)) void finish_rendezvous(int idx) {
    RankLds* L = ctx();
    Req* q = &reqs()[idx];
    rendezvous_copy((void*)(uintptr_t)q->buf, q->rv_src, q->got, q->rv_done, q->rv_val);
    q->state = RS_COMPLETE;
    L->pending_copies--;
}

// One pass over every active request in posting order: post waiting sends,
// collect rendezvous completions, match receives.  With `copies` (wait and
// test calls) a second pass then performs the copies of matched rendezvous
// receives, so every send this rank can post is out before it copies.
// A request freed by MPI_Request_free is released here once complete.
__device__ __forceinline__ bool reap_detached(int idx) {
    Req* q = &reqs()[idx];
    if (!(q->flags & RF_DETACHED) || q->state != RS_COMPLETE) return false;
    req_release(idx);
    ctx()->detached--;
    return true;
}

__device__ inline void progress(bool copies) {
    RankLds* L = ctx();
    const int R = L->R;
    uint64_t blocked = 0;  // destinations (hashed) whose earlier send found no slot
    for (int i = L->ring_head; i < L->ring_tail; i++) {
        int idx;
        if (!ring_live(i, &idx)) continue;
        Req* q = &reqs()[idx];
        if (L->detached && reap_detached(idx)) continue;
        if (q->kind == REQ_SEND) {
            if (q->state == RS_PENDING) {
                const uint64_t hb = 1ull << (q->peer & 63);
                if (blocked & hb) continue;
                if (try_post_send(idx)) L->pending_sends--;
                else blocked |= hb;
            }
            if (q->state == RS_WAIT_DONE) {
                const uint32_t v = uniform_u32(
                    ld_relaxed_agent(&L->done_words[(uint64_t)L->rank * R + idx]));
                if (v == q->stamp) {
                    fence_acquire_agent();
                    q->state = RS_COMPLETE;
                    if (q->flags & RF_STAGED) L->bounce_users--;
                }
            }
        } else if (q->kind == REQ_RECV && q->state == RS_PENDING) {
            Found f;
            if (!find_match(idx, &f)) continue;
            const int e = earlier_taker(i, f);
            if (e < 0) deliver(idx, f.src, f.seq, f.env, f.has_pre, f.pre);
            else try_match(e);  // takes its own oldest match: f or an older one
        }
    }
    if (!copies || L->pending_copies == 0) return;
    for (int i = L->ring_head; i < L->ring_tail; i++) {
        int idx;
        if (!ring_live(i, &idx)) continue;
        const Req* q = &reqs()[idx];
        if (q->kind == REQ_RECV && q->state == RS_MATCHED) {
            finish_rendezvous(idx);
            if (L->detached) reap_detached(idx);
        }
    }
}

// Start a receive of up to `bytes` into `buf` from world rank `src`
// (or -1 = any member of comm).
__device__ inline __attribute__((noinline)) int start_recv(void* buf, uint64_t bytes, int src, int tag, int ctxid, MPI_Comm comm) {
    RankLds* L = ctx();
    const int idx = req_alloc();
    if (idx < 0) return -1;
    Req* q = &reqs()[idx];
    q->kind = REQ_RECV;
    q->state = RS_PENDING;
    q->flags = 0;
    q->peer = src;
    q->tag = tag;
    q->ctx = ctxid;
    q->bytes = bytes;
    q->buf = (uint64_t)(uintptr_t)buf;
    q->comm = (uint64_t)(uintptr_t)comm;
    q->got = 0;
    q->st_error = MPI_SUCCESS;
    if (src == MPI_PROC_NULL) {
        q->state = RS_COMPLETE;
        q->st_source = MPI_PROC_NULL;
        q->st_tag = MPI_ANY_TAG;
        return idx;
    }
    // An earlier unmatched receive that could take the same message must get
    // the first chance at it (MPI ordering of posted receives).
    const bool earlier = L->unmatched_any > 0 || (src < 0 ? false : unmatched_src()[src] > 0) ||
                         (src < 0 && L->ring_tail - L->ring_head > 1);
    if (src < 0) L->unmatched_any++;
    else unmatched_src()[src] = (uint16_t)(unmatched_src()[src] + 1);
    if (earlier) progress(false);
    else try_match(idx);
    return idx;
}

// Status of a completed request (source translated to a rank of its comm).
// Every lane reads the request before any lane goes on to release it.
__device__ inline void fill_status(int idx, MPI_Status* st) {
    wave_sync();
    if (!st) return;
    const Req* q = &reqs()[idx];
    int src = q->st_source;
    if (q->kind == REQ_RECV && src >= 0 && (MPI_Comm)(uintptr_t)q->comm != MPI_COMM_WORLD)
        src = comm_rank_of_world(comm_get((MPI_Comm)(uintptr_t)q->comm), src);
    st->MPI_SOURCE = q->kind == REQ_RECV ? src : MPI_ANY_SOURCE;
    st->MPI_TAG = q->kind == REQ_RECV ? q->st_tag : MPI_ANY_TAG;
    st->MPI_ERROR = q->st_error;
    st->gpumpi_count_lo = (int)(uint32_t)q->got;
    st->gpumpi_count_hi = (int)(uint32_t)(q->got >> 32);
}

// Wait for request idx; fills the status and frees the request.
__device__ inline __attribute__((noinline)) int wait_req(int idx, MPI_Status* st) {
    RankLds* L = ctx();
    Req* q = &reqs()[idx];
    Spinner sp;
    // A receive that is the only unmatched one able to take messages from its
    // source polls that channel alone; the full progress pass (other requests
    // of this rank) runs every 16th poll.
    if (q->kind == REQ_RECV && q->state == RS_PENDING && q->peer >= 0 && L->unmatched_any == 0 &&
        unmatched_src()[q->peer] == 1) {
        int k = 0;
        while (q->state == RS_PENDING) {
            Found f;
            if (match_from(idx, q->peer, &f)) {
                deliver(idx, f.src, f.seq, f.env, f.has_pre, f.pre);
                break;
            }
            if ((++k & 15) == 0) progress();
            if (!sp.keep_polling()) break;
        }
    }
    while (q->state != RS_COMPLETE) {
        progress();
        if (q->state == RS_COMPLETE) break;
        if (!sp.keep_going()) {
            q->st_error = MPI_ERR_OTHER;
            break;
        }
    }
    fill_status(idx, st);
    const int err = q->st_error;
    req_release(idx);
    return err;
}

// Blocking eager send without a request: legal when no earlier send of this
// rank is waiting for a slot (the channel order is then the program order).
// Returns false when the message must take the general path.
__device__ inline bool send_direct(const void* buf, uint64_t n, int dst, int tag, int ctxid) {
    RankLds* L = ctx();
    if (L->pending_sends != 0 || n > (uint64_t)L->eager_bytes) return false;
    uint32_t s;
    if (!slot_reserve(dst, &s)) return false;
    slot_write_eager(dst, s, buf, n, ctxid, tag);
    return true;
}

// Blocking receive without a request: legal when no posted receive of this
// rank could take a message from `src` (no unmatched receive names `src` or
// MPI_ANY_SOURCE), so the oldest matching message of that channel is this
// receive's.  Polls the one channel; the full progress pass (the rank's
// other requests) runs every 16th poll.
__device__ inline int recv_direct(void* buf, uint64_t cap, int src, int tag, int ctxid, MPI_Comm comm,
                                  MPI_Status* st) {
    RankLds* L = ctx();
    Spinner sp;
    int k = 0;
    for (;;) {
        Found f;
        if (scan_channel(src, ctxid, tag, &f)) {
            Taken t;
            slot_take(f.src, f.seq, buf, cap, f.has_pre, f.pre, &t);
            if (t.rndv) rendezvous_copy(buf, t.rv_src, t.got, t.rv_done, t.rv_val);
            if (st) {
                st->MPI_SOURCE = comm == MPI_COMM_WORLD ? src : comm_rank_of_world(comm_get(comm), src);
                st->MPI_TAG = env_tag(f.env);
                st->MPI_ERROR = t.err;
                st->gpumpi_count_lo = (int)(uint32_t)t.got;
                st->gpumpi_count_hi = (int)(uint32_t)(t.got >> 32);
            }
            if (t.err != MPI_SUCCESS) GPUMPI_FAIL(t.err, "message truncated: receive buffer too small");
            return t.err;
        }
        // Other requests of this rank (queued sends, matched rendezvous
        // copies) progress every 16th poll; with none live there is nothing
        // to do.
        if ((++k & 15) == 0 && L->ring_tail != L->ring_head) progress();
        if (!sp.keep_polling()) return MPI_ERR_OTHER;
    }
}

}  // namespace rt
}  // namespace gpumpi
