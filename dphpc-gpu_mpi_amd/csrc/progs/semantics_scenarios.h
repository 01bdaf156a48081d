/*
 * Point-to-point, request and communicator semantics scenarios, written once
 * against the standard MPI C API.  The same text is compiled
 *   - as a device rank program of this runtime (progs/semantics.hip), and
 *   - as a host program under MPICH (oracle/hostmpi/semantics_golden.c),
 *     whose records are the golden fixtures the GPU tests compare with.
 *
 * Each scenario appends integers to its rank's record.  Only values that the
 * MPI standard fixes are recorded (payloads, per-source message order,
 * matching order of posted receives, status fields, counts, communicator
 * ranks and sizes); implementation constants (MPI_PROC_NULL, MPI_ANY_TAG,
 * MPI_UNDEFINED) are recorded as fixed stand-ins, and arrival order across
 * different sources, which MPI leaves open, is never recorded.
 *
 * The includer defines:
 *   SEM_FN           storage/target qualifier of the scenario functions
 *   SEM_ALLOC(b)     b bytes of memory every rank's peers may read (heap)
 *   SEM_FREE(p, b)   release it
 *
 * What is exercised, and where the reference stood (gpu_libs/gpu_mpi/mpi.cu):
 *   Sendrecv (a stub there, :255-260), ANY_SOURCE / ANY_TAG receives against
 *   queued and later messages (cuda_mpi.cu:296-310 never matched queued ones),
 *   the order of posted receives, Test / Testall / Waitany / Waitsome
 *   (:573-597), Request_free (:608-617), Ssend / Issend, Status and
 *   Get_count (:605 left Status unfilled), PROC_NULL, Group_incl /
 *   Group_range_incl / Comm_create / Comm_split with MPI_UNDEFINED
 *   (communicator.cu:74-213, group.cu:53-122).
 */
#ifndef SEM_FN
#error "define SEM_FN before including semantics_scenarios.h"
#endif

#define SEM_PROC_NULL_MARK (-100)
#define SEM_ANY_TAG_MARK (-101)
#define SEM_UNDEFINED_MARK (-102)
#define SEM_NULL_COMM_MARK (-103)
#define SEM_ANY_SOURCE_MARK (-104)

struct sem_rec {
    int* v;
    int n;
    int cap;
};

SEM_FN void sem_put(struct sem_rec* r, int x) {
    if (r->n < r->cap) r->v[r->n] = x;
    r->n++;
}
SEM_FN int sem_src(int s) {
    return s == MPI_PROC_NULL ? SEM_PROC_NULL_MARK : s == MPI_ANY_SOURCE ? SEM_ANY_SOURCE_MARK : s;
}
SEM_FN int sem_tag(int t) { return t == MPI_ANY_TAG ? SEM_ANY_TAG_MARK : t; }
SEM_FN int sem_cnt(int c) { return c == MPI_UNDEFINED ? SEM_UNDEFINED_MARK : c; }
/* Order-sensitive checksum of n ints (unsigned wraparound). */
SEM_FN int sem_sum(const int* a, int n) {
    unsigned h = 2166136261u;
    int i;
    for (i = 0; i < n; i++) h = (h ^ (unsigned)a[i]) * 16777619u;
    return (int)(h & 0x7fffffffu);
}
SEM_FN void sem_status(struct sem_rec* r, const MPI_Status* st, MPI_Datatype dt) {
    int c = 0;
    MPI_Get_count(st, dt, &c);
    sem_put(r, sem_src(st->MPI_SOURCE));
    sem_put(r, sem_tag(st->MPI_TAG));
    sem_put(r, sem_cnt(c));
}

/* 1. Sendrecv around a ring (eager and rendezvous sizes), then along an open
 * chain whose ends name MPI_PROC_NULL. */
SEM_FN void sem_sendrecv(struct sem_rec* r, int rank, int P) {
    const int sizes[2] = {4, 20000};
    int k, i;
    for (k = 0; k < 2; k++) {
        const int n = sizes[k];
        int* sb = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
        int* rb = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
        MPI_Status st;
        for (i = 0; i < n; i++) {
            sb[i] = rank * 1000003 + i;
            rb[i] = -1;
        }
        MPI_Sendrecv(sb, n, MPI_INT, (rank + 1) % P, 40 + k, rb, n, MPI_INT, (rank + P - 1) % P, 40 + k,
                     MPI_COMM_WORLD, &st);
        sem_put(r, sem_sum(rb, n));
        sem_status(r, &st, MPI_INT);
        SEM_FREE(rb, sizeof(int) * (size_t)n);
        SEM_FREE(sb, sizeof(int) * (size_t)n);
    }
    {
        int* x = (int*)SEM_ALLOC(sizeof(int) * 8);
        int* y = (int*)SEM_ALLOC(sizeof(int) * 8);
        MPI_Status st;
        const int up = rank + 1 < P ? rank + 1 : MPI_PROC_NULL;
        const int down = rank > 0 ? rank - 1 : MPI_PROC_NULL;
        for (i = 0; i < 8; i++) {
            x[i] = 100 * rank + i;
            y[i] = -7;
        }
        MPI_Sendrecv(x, 8, MPI_INT, up, 44, y, 8, MPI_INT, down, 44, MPI_COMM_WORLD, &st);
        sem_put(r, sem_sum(y, 8));
        sem_status(r, &st, MPI_INT);
        SEM_FREE(y, sizeof(int) * 8);
        SEM_FREE(x, sizeof(int) * 8);
    }
}

/* 2. MPI_ANY_SOURCE / MPI_ANY_TAG receives at rank 0 of messages already
 * queued (sent before a barrier), then of messages that arrive after the
 * receives were posted.  Recorded per source, in the order received from
 * that source: MPI's non-overtaking rule fixes it. */
SEM_FN void sem_any_source(struct sem_rec* r, int rank, int P) {
    const int K = 3;
    int i, j, s;
    if (rank > 0)
        for (j = 0; j < K; j++) {
            int v = rank * 100 + j;
            MPI_Send(&v, 1, MPI_INT, 0, 10 + j, MPI_COMM_WORLD);
        }
    MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 0) {
        int* seen = (int*)SEM_ALLOC(sizeof(int) * (size_t)(P * K));
        int* cnt = (int*)SEM_ALLOC(sizeof(int) * (size_t)P);
        int* v = (int*)SEM_ALLOC(sizeof(int));
        int bad = 0;
        for (s = 0; s < P; s++) cnt[s] = 0;
        for (i = 0; i < (P - 1) * K; i++) {
            MPI_Status st;
            *v = -1;
            MPI_Recv(v, 1, MPI_INT, MPI_ANY_SOURCE, MPI_ANY_TAG, MPI_COMM_WORLD, &st);
            s = st.MPI_SOURCE;
            if (s <= 0 || s >= P || *v / 100 != s || st.MPI_TAG != 10 + *v % 100) {
                bad++;
                continue;
            }
            seen[s * K + cnt[s]] = *v;
            cnt[s]++;
        }
        sem_put(r, bad);
        for (s = 1; s < P; s++)
            for (j = 0; j < K; j++) sem_put(r, j < cnt[s] ? seen[s * K + j] : -1);
        SEM_FREE(v, sizeof(int));
        SEM_FREE(cnt, sizeof(int) * (size_t)P);
        SEM_FREE(seen, sizeof(int) * (size_t)(P * K));
    }
    MPI_Barrier(MPI_COMM_WORLD);
    /* Receives posted first; the messages arrive later.  Request i is posted
     * before request i+1, so the k-th message of a source lands in an
     * earlier request than its (k+1)-th. */
    if (rank == 0) {
        const int n = (P - 1) * K;
        MPI_Request* rq = (MPI_Request*)SEM_ALLOC(sizeof(MPI_Request) * (size_t)n);
        int* buf = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
        int* nxt = (int*)SEM_ALLOC(sizeof(int) * (size_t)P);
        int* seen = (int*)SEM_ALLOC(sizeof(int) * (size_t)(P * K));
        int bad = 0;
        for (i = 0; i < n; i++) MPI_Irecv(&buf[i], 1, MPI_INT, MPI_ANY_SOURCE, MPI_ANY_TAG, MPI_COMM_WORLD, &rq[i]);
        MPI_Barrier(MPI_COMM_WORLD);
        MPI_Waitall(n, rq, MPI_STATUSES_IGNORE);
        for (s = 0; s < P; s++) nxt[s] = 0;
        for (i = 0; i < n; i++) {
            s = buf[i] / 100;
            if (s <= 0 || s >= P || nxt[s] >= K) {
                bad++;
                continue;
            }
            seen[s * K + nxt[s]] = buf[i];
            nxt[s]++;
        }
        sem_put(r, bad);
        for (s = 1; s < P; s++)
            for (j = 0; j < K; j++) sem_put(r, j < nxt[s] ? seen[s * K + j] : -1);
        SEM_FREE(seen, sizeof(int) * (size_t)(P * K));
        SEM_FREE(nxt, sizeof(int) * (size_t)P);
        SEM_FREE(buf, sizeof(int) * (size_t)n);
        SEM_FREE(rq, sizeof(MPI_Request) * (size_t)n);
    } else {
        MPI_Barrier(MPI_COMM_WORLD);
        for (j = 0; j < K; j++) {
            int v = rank * 100 + j;
            MPI_Send(&v, 1, MPI_INT, 0, 20 + j, MPI_COMM_WORLD);
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
}

/* 3. ANY_TAG from one source takes messages in send order; specific tags
 * can be received out of send order. */
SEM_FN void sem_any_tag(struct sem_rec* r, int rank, int P) {
    int i;
    (void)P;
    if (rank == 1) {
        const int tags[3] = {5, 3, 9};
        for (i = 0; i < 3; i++) {
            int v = tags[i] * 10;
            MPI_Send(&v, 1, MPI_INT, 0, tags[i], MPI_COMM_WORLD);
        }
        for (i = 1; i <= 2; i++) {
            int v = 11 * i;
            MPI_Send(&v, 1, MPI_INT, 0, i, MPI_COMM_WORLD);
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 0) {
        for (i = 0; i < 3; i++) {
            int v = -1;
            MPI_Status st;
            MPI_Recv(&v, 1, MPI_INT, 1, MPI_ANY_TAG, MPI_COMM_WORLD, &st);
            sem_put(r, v);
            sem_status(r, &st, MPI_INT);
        }
        for (i = 2; i >= 1; i--) {
            int v = -1;
            MPI_Status st;
            MPI_Recv(&v, 1, MPI_INT, 1, i, MPI_COMM_WORLD, &st);
            sem_put(r, v);
            sem_put(r, st.MPI_TAG);
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
}

/* 4. Posted-receive order: a message that several posted receives accept goes
 * to the earliest posted one, including a wildcard posted before a specific
 * receive and a specific receive posted before a wildcard. */
SEM_FN void sem_posted_order(struct sem_rec* r, int rank, int P) {
    if (rank == 0) {
        int* v = (int*)SEM_ALLOC(sizeof(int) * 4); /* a, b, c, d */
        MPI_Request rq[4];
        MPI_Status st[4];
        int i;
        for (i = 0; i < 4; i++) v[i] = -1;
        MPI_Irecv(&v[0], 1, MPI_INT, MPI_ANY_SOURCE, 7, MPI_COMM_WORLD, &rq[0]);
        MPI_Irecv(&v[1], 1, MPI_INT, 1, 7, MPI_COMM_WORLD, &rq[1]);
        MPI_Irecv(&v[2], 1, MPI_INT, 2 % P, MPI_ANY_TAG, MPI_COMM_WORLD, &rq[2]);
        MPI_Irecv(&v[3], 1, MPI_INT, MPI_ANY_SOURCE, 8, MPI_COMM_WORLD, &rq[3]);
        MPI_Barrier(MPI_COMM_WORLD);
        MPI_Waitall(4, rq, st);
        for (i = 0; i < 4; i++) sem_put(r, v[i]);
        for (i = 0; i < 4; i++) {
            sem_put(r, sem_src(st[i].MPI_SOURCE));
            sem_put(r, sem_tag(st[i].MPI_TAG));
        }
        SEM_FREE(v, sizeof(int) * 4);
    } else {
        /* Whatever the arrival order: 71 precedes 72 (same source), both
         * reach the earliest posted receive that accepts them; 81 is
         * accepted by c and d and goes to c, posted first; 82 only by d. */
        MPI_Barrier(MPI_COMM_WORLD);
        if (rank == 1) {
            int v = 71;
            MPI_Send(&v, 1, MPI_INT, 0, 7, MPI_COMM_WORLD);
            v = 72;
            MPI_Send(&v, 1, MPI_INT, 0, 7, MPI_COMM_WORLD);
        }
        if (rank == 2) {
            int v = 81;
            MPI_Send(&v, 1, MPI_INT, 0, 8, MPI_COMM_WORLD);
        }
        if (rank == 3) {
            int v = 82;
            MPI_Send(&v, 1, MPI_INT, 0, 8, MPI_COMM_WORLD);
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
}

/* 5. Test polling, Testall, Waitany and Waitsome at rank 0 over one receive
 * from every other rank; results are recorded by request index. */
SEM_FN void sem_completion(struct sem_rec* r, int rank, int P) {
    int i, round;
    if (rank == 0) {
        MPI_Request q = MPI_REQUEST_NULL;
        MPI_Status st;
        int flag = 0;
        int* v = (int*)SEM_ALLOC(sizeof(int));
        long polls = 0;
        *v = -1;
        MPI_Test(&q, &flag, &st);
        sem_put(r, flag);
        sem_put(r, sem_src(st.MPI_SOURCE));
        MPI_Irecv(v, 1, MPI_INT, 1, 50, MPI_COMM_WORLD, &q);
        MPI_Barrier(MPI_COMM_WORLD);
        flag = 0;
        while (!flag && polls < 100000000L) {
            MPI_Test(&q, &flag, &st);
            polls++;
        }
        sem_put(r, flag);
        sem_put(r, *v);
        sem_status(r, &st, MPI_INT);
        sem_put(r, q == MPI_REQUEST_NULL);
        SEM_FREE(v, sizeof(int));
    } else {
        MPI_Barrier(MPI_COMM_WORLD);
        if (rank == 1) {
            int v = 5050;
            MPI_Send(&v, 1, MPI_INT, 0, 50, MPI_COMM_WORLD);
        }
    }
    for (round = 0; round < 3; round++) {
        const int tag = 60 + round;
        if (rank == 0) {
            const int n = P - 1;
            MPI_Request* rq = (MPI_Request*)SEM_ALLOC(sizeof(MPI_Request) * (size_t)n);
            MPI_Status* sts = (MPI_Status*)SEM_ALLOC(sizeof(MPI_Status) * (size_t)n);
            int* idx = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
            int* val = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
            int* done = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
            int completions = 0, calls = 0, bad = 0;
            for (i = 0; i < n; i++) {
                val[i] = -1;
                done[i] = 0;
                MPI_Irecv(&val[i], 1, MPI_INT, i + 1, tag, MPI_COMM_WORLD, &rq[i]);
            }
            MPI_Barrier(MPI_COMM_WORLD);
            if (round == 0) { /* Waitany until every request is null */
                for (;;) {
                    int k = -5;
                    MPI_Status st;
                    MPI_Waitany(n, rq, &k, &st);
                    if (k == MPI_UNDEFINED) break;
                    if (k < 0 || k >= n || done[k] || st.MPI_SOURCE != k + 1) bad++;
                    else done[k] = 1;
                    completions++;
                    if (++calls > 4 * n + 4) break;
                }
            } else if (round == 1) { /* Waitsome until MPI_UNDEFINED */
                for (;;) {
                    int oc = -5;
                    MPI_Waitsome(n, rq, &oc, idx, sts);
                    if (oc == MPI_UNDEFINED) break;
                    for (i = 0; i < oc; i++) {
                        const int k = idx[i];
                        if (k < 0 || k >= n || done[k] || sts[i].MPI_SOURCE != k + 1) bad++;
                        else done[k] = 1;
                    }
                    completions += oc;
                    if (++calls > 4 * n + 4) break;
                }
            } else { /* Testall until true */
                int flag = 0;
                long polls = 0;
                while (!flag && polls < 100000000L) {
                    MPI_Testall(n, rq, &flag, sts);
                    polls++;
                }
                completions = flag ? n : 0;
                for (i = 0; i < n && flag; i++) {
                    if (sts[i].MPI_SOURCE != i + 1 || rq[i] != MPI_REQUEST_NULL) bad++;
                    done[i] = 1;
                }
            }
            sem_put(r, completions);
            sem_put(r, bad);
            for (i = 0; i < n; i++) sem_put(r, done[i] ? val[i] : -1);
            SEM_FREE(done, sizeof(int) * (size_t)n);
            SEM_FREE(val, sizeof(int) * (size_t)n);
            SEM_FREE(idx, sizeof(int) * (size_t)n);
            SEM_FREE(sts, sizeof(MPI_Status) * (size_t)n);
            SEM_FREE(rq, sizeof(MPI_Request) * (size_t)n);
        } else {
            int v = 1000 * (round + 1) + rank;
            MPI_Barrier(MPI_COMM_WORLD);
            MPI_Send(&v, 1, MPI_INT, 0, tag, MPI_COMM_WORLD);
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
}

/* 6. MPI_Request_free is local: rank 0 frees the handle of a rendezvous-size
 * Isend and then receives from rank 1, which first sends to rank 0 and only
 * then receives rank 0's message.  A Request_free that waited for its send
 * would deadlock here. */
SEM_FN void sem_request_free(struct sem_rec* r, int rank, int P) {
    const int n = 4096;
    int i;
    (void)P;
    if (rank <= 1) {
        int* sb = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
        int* rb = (int*)SEM_ALLOC(sizeof(int) * (size_t)n);
        for (i = 0; i < n; i++) {
            sb[i] = (rank + 1) * 7919 + 3 * i;
            rb[i] = 0;
        }
        if (rank == 0) {
            MPI_Request q;
            MPI_Isend(sb, n, MPI_INT, 1, 3, MPI_COMM_WORLD, &q);
            MPI_Request_free(&q);
            sem_put(r, q == MPI_REQUEST_NULL);
            MPI_Recv(rb, n, MPI_INT, 1, 4, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        } else {
            MPI_Send(sb, n, MPI_INT, 0, 4, MPI_COMM_WORLD);
            MPI_Recv(rb, n, MPI_INT, 0, 3, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        }
        sem_put(r, sem_sum(rb, n));
        /* Free of a request that has already completed. */
        if (rank == 1) {
            int v = 99;
            MPI_Request q;
            MPI_Isend(&v, 1, MPI_INT, 0, 5, MPI_COMM_WORLD, &q);
            MPI_Request_free(&q);
        } else {
            int v = -1;
            MPI_Recv(&v, 1, MPI_INT, 1, 5, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
            sem_put(r, v);
        }
        MPI_Barrier(MPI_COMM_WORLD);
        SEM_FREE(rb, sizeof(int) * (size_t)n);
        SEM_FREE(sb, sizeof(int) * (size_t)n);
    } else {
        MPI_Barrier(MPI_COMM_WORLD);
    }
}

/* 7. Synchronous sends (eager-size and rendezvous-size), Get_count for
 * whole and partial elements, and receives from MPI_PROC_NULL. */
SEM_FN void sem_sync_status(struct sem_rec* r, int rank, int P) {
    int i;
    (void)P;
    if (rank == 0) {
        int x[16];
        int* big = (int*)SEM_ALLOC(sizeof(int) * 3000);
        MPI_Request q;
        for (i = 0; i < 16; i++) x[i] = 16 - i;
        for (i = 0; i < 3000; i++) big[i] = i * i;
        MPI_Ssend(x, 16, MPI_INT, 1, 70, MPI_COMM_WORLD);
        MPI_Issend(big, 3000, MPI_INT, 1, 71, MPI_COMM_WORLD, &q);
        MPI_Wait(&q, MPI_STATUS_IGNORE);
        MPI_Send(x, 3, MPI_INT, 1, 72, MPI_COMM_WORLD); /* 12 bytes */
        SEM_FREE(big, sizeof(int) * 3000);
    } else if (rank == 1) {
        int y[100];
        int* big = (int*)SEM_ALLOC(sizeof(int) * 4000);
        MPI_Status st;
        int c = 0;
        MPI_Recv(y, 100, MPI_INT, 0, 70, MPI_COMM_WORLD, &st);
        sem_put(r, sem_sum(y, 16));
        sem_status(r, &st, MPI_INT);
        MPI_Get_count(&st, MPI_DOUBLE, &c);
        sem_put(r, sem_cnt(c));
        MPI_Recv(big, 4000, MPI_INT, 0, 71, MPI_COMM_WORLD, &st);
        sem_put(r, sem_sum(big, 3000));
        sem_status(r, &st, MPI_INT);
        MPI_Recv(y, 100, MPI_INT, 0, 72, MPI_COMM_WORLD, &st);
        sem_status(r, &st, MPI_INT);
        MPI_Get_count(&st, MPI_DOUBLE, &c); /* 12 bytes: not whole doubles */
        sem_put(r, sem_cnt(c));
        MPI_Get_count(&st, MPI_SHORT, &c);
        sem_put(r, sem_cnt(c));
        SEM_FREE(big, sizeof(int) * 4000);
    }
    {
        int v = 123;
        MPI_Status st;
        MPI_Send(&v, 1, MPI_INT, MPI_PROC_NULL, 1, MPI_COMM_WORLD);
        MPI_Recv(&v, 1, MPI_INT, MPI_PROC_NULL, 1, MPI_COMM_WORLD, &st);
        sem_put(r, v);
        sem_status(r, &st, MPI_INT);
    }
    MPI_Barrier(MPI_COMM_WORLD);
}

/* 8. Groups and communicators: Group_incl of the even ranks, Group_range_incl
 * of a strided descending range, Comm_create on both, Comm_split with
 * MPI_UNDEFINED, collectives and p2p on the results, Comm_free. */
SEM_FN void sem_groups(struct sem_rec* r, int rank, int P) {
    MPI_Group world, evens, rng;
    MPI_Comm ce, cr, cs;
    int* ranks = (int*)SEM_ALLOC(sizeof(int) * (size_t)P);
    int ne = 0, i, sz = -1, me = -1, sum = -1;
    int range[1][3];
    MPI_Comm_group(MPI_COMM_WORLD, &world);
    for (i = 0; i < P; i += 2) ranks[ne++] = i;
    MPI_Group_incl(world, ne, ranks, &evens);
    MPI_Group_size(evens, &sz);
    MPI_Group_rank(evens, &me);
    sem_put(r, sz);
    sem_put(r, sem_cnt(me));
    MPI_Comm_create(MPI_COMM_WORLD, evens, &ce);
    if (ce == MPI_COMM_NULL) {
        sem_put(r, SEM_NULL_COMM_MARK);
    } else {
        int v = rank;
        MPI_Comm_rank(ce, &me);
        MPI_Comm_size(ce, &sz);
        MPI_Allreduce(&v, &sum, 1, MPI_INT, MPI_SUM, ce);
        sem_put(r, me);
        sem_put(r, sz);
        sem_put(r, sum);
        MPI_Comm_free(&ce);
        sem_put(r, ce == MPI_COMM_NULL);
    }
    range[0][0] = P - 1;
    range[0][1] = 0;
    range[0][2] = -3;
    MPI_Group_range_incl(world, 1, range, &rng);
    MPI_Group_size(rng, &sz);
    MPI_Group_rank(rng, &me);
    sem_put(r, sz);
    sem_put(r, sem_cnt(me));
    MPI_Comm_create(MPI_COMM_WORLD, rng, &cr);
    if (cr == MPI_COMM_NULL) {
        sem_put(r, SEM_NULL_COMM_MARK);
    } else {
        int mx = -1, v, n2;
        MPI_Comm_rank(cr, &me);
        MPI_Comm_size(cr, &n2);
        v = rank * 10 + me;
        MPI_Allreduce(&v, &mx, 1, MPI_INT, MPI_MAX, cr);
        sem_put(r, me);
        sem_put(r, n2);
        sem_put(r, mx);
        /* A ring shift inside the new communicator (ranks translate). */
        {
            int* got = (int*)SEM_ALLOC(sizeof(int));
            MPI_Status st;
            *got = -1;
            MPI_Sendrecv(&rank, 1, MPI_INT, (me + 1) % n2, 9, got, 1, MPI_INT, (me + n2 - 1) % n2, 9, cr, &st);
            sem_put(r, *got);
            sem_put(r, st.MPI_SOURCE);
            SEM_FREE(got, sizeof(int));
        }
        MPI_Comm_free(&cr);
    }
    /* Split: ranks divisible by 3 get MPI_UNDEFINED. */
    MPI_Comm_split(MPI_COMM_WORLD, rank % 3 == 0 ? MPI_UNDEFINED : rank % 3, P - rank, &cs);
    if (cs == MPI_COMM_NULL) {
        sem_put(r, SEM_NULL_COMM_MARK);
    } else {
        MPI_Comm_rank(cs, &me);
        MPI_Comm_size(cs, &sz);
        sem_put(r, me);
        sem_put(r, sz);
        MPI_Comm_free(&cs);
    }
    MPI_Group_free(&rng);
    MPI_Group_free(&evens);
    SEM_FREE(ranks, sizeof(int) * (size_t)P);
    MPI_Barrier(MPI_COMM_WORLD);
}

/* Scenario table: bit k of the mask runs scenario k (all need P >= 4). */
#define SEM_NUM_SCENARIOS 8
SEM_FN void sem_run(struct sem_rec* r, int rank, int P, int mask) {
    if (mask & 1) sem_sendrecv(r, rank, P);
    if (mask & 2) sem_any_source(r, rank, P);
    if (mask & 4) sem_any_tag(r, rank, P);
    if (mask & 8) sem_posted_order(r, rank, P);
    if (mask & 16) sem_completion(r, rank, P);
    if (mask & 32) sem_request_free(r, rank, P);
    if (mask & 64) sem_sync_status(r, rank, P);
    if (mask & 128) sem_groups(r, rank, P);
}
