// Definitions of the device MPI_* API declared in include/gpumpi/mpi.h.
// Each entry point cites the reference function whose contract it keeps.
#pragma once

#include "coll.hpp"

namespace gpumpi {
namespace rt {

__device__ inline int msg_bytes(int count, MPI_Datatype dt, uint64_t* out) {
    const int es = type_size(dt);
    if (es <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype");
    if (count < 0) return GPUMPI_FAIL(MPI_ERR_COUNT, "negative count");
    *out = (uint64_t)count * (uint64_t)es;
    return MPI_SUCCESS;
}

// Resolve (comm, rank) of a p2p call to a world rank and the communicator's
// context id; MPI_PROC_NULL and MPI_ANY_SOURCE pass through.  WORLD (context
// 1, identity rank map) needs no lookup.
__device__ inline int p2p_peer(MPI_Comm comm, int rank, bool recv, int* ctx_out, int* wout) {
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "point-to-point on MPI_COMM_NULL");
    const bool pass = rank == MPI_PROC_NULL || (recv && rank == MPI_ANY_SOURCE);
    if (comm == MPI_COMM_WORLD) {
        *ctx_out = 1;
        if (!pass && (rank < 0 || rank >= ctx()->nranks))
            return GPUMPI_FAIL(MPI_ERR_RANK, "invalid rank in point-to-point call");
        *wout = rank;
        return MPI_SUCCESS;
    }
    CommShared* c = comm_get(comm);
    *ctx_out = c->ctx;
    if (pass) {
        *wout = rank;
        return MPI_SUCCESS;
    }
    if (rank < 0 || rank >= c->size) return GPUMPI_FAIL(MPI_ERR_RANK, "invalid rank in point-to-point call");
    *wout = comm_world_of(c, rank);
    return MPI_SUCCESS;
}

__device__ inline int check_tag(int tag, bool recv) {
    if (tag > MPI_TAG_UB || (tag < 0 && !(recv && tag == MPI_ANY_TAG)))
        return GPUMPI_FAIL(MPI_ERR_TAG, "invalid tag");
    return MPI_SUCCESS;
}

__device__ inline int isend_common(const void* buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm,
                                   bool sync, int* idx_out) {
    uint64_t bytes;
    int err = msg_bytes(count, dt, &bytes);
    if (err) return err;
    if ((err = check_tag(tag, false))) return err;
    int cid, wdst;
    if ((err = p2p_peer(comm, dest, false, &cid, &wdst))) return err;
    const int idx = start_send(buf, bytes, wdst, tag, cid, comm, sync);
    if (idx < 0) return MPI_ERR_REQUEST;
    *idx_out = idx;
    return MPI_SUCCESS;
}

__device__ inline int irecv_common(void* buf, int count, MPI_Datatype dt, int source, int tag, MPI_Comm comm,
                                   int* idx_out) {
    uint64_t bytes;
    int err = msg_bytes(count, dt, &bytes);
    if (err) return err;
    if ((err = check_tag(tag, true))) return err;
    int cid, wsrc;
    if ((err = p2p_peer(comm, source, true, &cid, &wsrc))) return err;
    const int idx = start_recv(buf, bytes, wsrc, tag, cid, comm);
    if (idx < 0) return MPI_ERR_REQUEST;
    *idx_out = idx;
    return MPI_SUCCESS;
}

}  // namespace rt
}  // namespace gpumpi

// ---------------------------------------------------------------------------
// Environment (mpi.cu:44-75, 142-146, 271-274)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Init(int*, char***) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    ctx()->initialized = 1;
    return comm_barrier(world()->comm_world, ctx()->nranks) ? MPI_SUCCESS : MPI_ERR_OTHER;
}
__device__ inline int MPI_Init_thread(int* argc, char*** argv, int, int* provided) {
    if (provided) *provided = MPI_THREAD_FUNNELED;
    return MPI_Init(argc, argv);
}
__device__ inline int MPI_Initialized(int* flag) {
    *flag = gpumpi::rt::ctx()->initialized;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Finalize(void) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    // Requests whose handles were freed still have to complete.
    Spinner sp;
    while (ctx()->detached > 0) {
        progress();
        if (ctx()->detached == 0 || !sp.keep_going()) break;
    }
    const bool ok = comm_barrier(world()->comm_world, ctx()->nranks);
    ctx()->finalized = 1;
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}
__device__ inline int MPI_Abort(MPI_Comm, int errorcode) {
    gpumpi::rt::raise_error(errorcode ? errorcode : MPI_ERR_OTHER, __LINE__, "MPI_Abort called");
    return MPI_SUCCESS;
}
// mpi.cu:70-75: name "GPU thread" there; here the rank is a workgroup.
__device__ inline int MPI_Get_processor_name(char* name, int* resultlen) {
    const char n[] = "MI355X workgroup";
    for (int i = 0; i < (int)sizeof(n); i++) name[i] = n[i];
    *resultlen = (int)sizeof(n) - 1;
    return MPI_SUCCESS;
}
__device__ inline double MPI_Wtick(void) { return 1.0 / (double)gpumpi::rt::world()->cfg.wall_hz; }
__device__ inline double MPI_Wtime(void) {
    return (double)gpumpi::rt::uniform_u64(gpumpi::rt::wall_ticks()) * MPI_Wtick();
}

// ---------------------------------------------------------------------------
// Datatypes (mpi.cu:237-245, 283-290; datatypes.cu:10-15)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Type_size(MPI_Datatype datatype, int* size) {
    const int s = gpumpi::rt::type_size(datatype);
    if (s <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype");
    *size = s;
    return MPI_SUCCESS;
}
// MPI_Type_contiguous: `count` copies of oldtype, itself basic or contiguous
// (the reference's was NOT_IMPLEMENTED, datatypes.cu:10-15).  The handle
// encodes the basic type and the element count, so it needs no storage,
// commit or free work, and is the same value on every rank.
__device__ inline int MPI_Type_contiguous(int count, MPI_Datatype oldtype, MPI_Datatype* newtype) {
    using namespace gpumpi::rt;
    if (!newtype) return GPUMPI_FAIL(MPI_ERR_ARG, "MPI_Type_contiguous: null newtype");
    if (type_size(oldtype) <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "MPI_Type_contiguous: invalid oldtype");
    if (count < 1) return GPUMPI_FAIL(MPI_ERR_COUNT, "MPI_Type_contiguous: count must be at least 1");
    const int64_t n = (int64_t)count * type_count(oldtype);
    const int base = type_base(oldtype);
    if (n > kContigMaxCount || n * basic_type_size(base) > 0x7fffffff)
        return GPUMPI_FAIL(MPI_ERR_COUNT, "MPI_Type_contiguous: type too large");
    *newtype = (MPI_Datatype)(GPUMPI_CONTIGUOUS_FLAG | (int)(n << kContigShift) | base);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Type_commit(MPI_Datatype* datatype) {
    if (!datatype || gpumpi::rt::type_size(*datatype) <= 0)
        return GPUMPI_FAIL(MPI_ERR_TYPE, "MPI_Type_commit: invalid datatype");
    return MPI_SUCCESS;
}
__device__ inline int MPI_Type_free(MPI_Datatype* datatype) {
    if (!datatype || !gpumpi::rt::type_is_contig(*datatype))
        return GPUMPI_FAIL(MPI_ERR_TYPE, "MPI_Type_free: not a derived datatype");
    *datatype = MPI_DATATYPE_NULL;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Type_vector(int, int, int, MPI_Datatype, MPI_Datatype*) {
    return GPUMPI_FAIL(MPI_ERR_TYPE, "strided derived datatypes (MPI_Type_vector) are not supported");
}
__device__ inline int MPI_Get_count(const MPI_Status* st, MPI_Datatype dt, int* count) {
    const int s = gpumpi::rt::type_size(dt);
    if (s <= 0) return GPUMPI_FAIL(MPI_ERR_TYPE, "invalid datatype");
    const uint64_t b = (uint64_t)(uint32_t)st->gpumpi_count_lo | ((uint64_t)(uint32_t)st->gpumpi_count_hi << 32);
    *count = (b % (uint64_t)s) ? MPI_UNDEFINED : (int)(b / (uint64_t)s);
    return MPI_SUCCESS;
}

// ---------------------------------------------------------------------------
// Operators (operators.cu:241-254)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Op_create(MPI_User_function* fn, int commute, MPI_Op* op) {
    MPI_Op_impl* o = (MPI_Op_impl*)gpumpi::rt::rank_malloc(sizeof(MPI_Op_impl));
    if (!o) return MPI_ERR_NO_MEM;
    o->fn = fn;
    o->commute = commute;
    *op = o;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Op_free(MPI_Op* op) {
    if ((uintptr_t)*op > GPUMPI_NUM_BUILTIN_OPS) gpumpi::rt::rank_free(*op, sizeof(MPI_Op_impl));
    *op = MPI_OP_NULL;
    return MPI_SUCCESS;
}

// ---------------------------------------------------------------------------
// Point to point (mpi.cu:247-269, 543-617; cuda_mpi.cu:65-471)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Isend(const void* buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm,
                                MPI_Request* request) {
    gpumpi::rt::WaveJoin join_;
    int idx;
    int err = gpumpi::rt::isend_common(buf, count, dt, dest, tag, comm, false, &idx);
    if (err) return err;
    *request = gpumpi::rt::req_handle(idx);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Issend(const void* buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm,
                                 MPI_Request* request) {
    gpumpi::rt::WaveJoin join_;
    int idx;
    int err = gpumpi::rt::isend_common(buf, count, dt, dest, tag, comm, true, &idx);
    if (err) return err;
    *request = gpumpi::rt::req_handle(idx);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Irecv(void* buf, int count, MPI_Datatype dt, int source, int tag, MPI_Comm comm,
                                MPI_Request* request) {
    gpumpi::rt::WaveJoin join_;
    int idx;
    int err = gpumpi::rt::irecv_common(buf, count, dt, source, tag, comm, &idx);
    if (err) return err;
    *request = gpumpi::rt::req_handle(idx);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Send(const void* buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    uint64_t bytes;
    int err = msg_bytes(count, dt, &bytes);
    if (err) return err;
    if ((err = check_tag(tag, false))) return err;
    int cid, wdst;
    if ((err = p2p_peer(comm, dest, false, &cid, &wdst))) return err;
    if (wdst == MPI_PROC_NULL) return MPI_SUCCESS;
    if (send_direct(buf, bytes, wdst, tag, cid)) return MPI_SUCCESS;
    const int idx = start_send(buf, bytes, wdst, tag, cid, comm, false);
    if (idx < 0) return MPI_ERR_REQUEST;
    return wait_req(idx, MPI_STATUS_IGNORE);
}
__device__ inline int MPI_Ssend(const void* buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm) {
    gpumpi::rt::WaveJoin join_;
    int idx;
    int err = gpumpi::rt::isend_common(buf, count, dt, dest, tag, comm, true, &idx);
    if (err) return err;
    return gpumpi::rt::wait_req(idx, MPI_STATUS_IGNORE);
}
__device__ inline int MPI_Recv(void* buf, int count, MPI_Datatype dt, int source, int tag, MPI_Comm comm,
                               MPI_Status* status) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    uint64_t bytes;
    int err = msg_bytes(count, dt, &bytes);
    if (err) return err;
    if ((err = check_tag(tag, true))) return err;
    int cid, wsrc;
    if ((err = p2p_peer(comm, source, true, &cid, &wsrc))) return err;
    if (wsrc >= 0 && ctx()->unmatched_any == 0 && unmatched_src()[wsrc] == 0)
        return recv_direct(buf, bytes, wsrc, tag, cid, comm, status);
    const int idx = start_recv(buf, bytes, wsrc, tag, cid, comm);
    if (idx < 0) return MPI_ERR_REQUEST;
    return wait_req(idx, status);
}
// A real Sendrecv (the reference's was a stub, mpi.cu:255-260): both halves
// are posted before either is waited on, so exchanges cannot deadlock.
__device__ inline int MPI_Sendrecv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, int dest,
                                   int sendtag, void* recvbuf, int recvcount, MPI_Datatype recvtype, int source,
                                   int recvtag, MPI_Comm comm, MPI_Status* status) {
    gpumpi::rt::WaveJoin join_;
    int ri, si;
    int err = gpumpi::rt::irecv_common(recvbuf, recvcount, recvtype, source, recvtag, comm, &ri);
    if (err) return err;
    err = gpumpi::rt::isend_common(sendbuf, sendcount, sendtype, dest, sendtag, comm, false, &si);
    if (err) return err;
    int e1 = gpumpi::rt::wait_req(si, MPI_STATUS_IGNORE);
    int e2 = gpumpi::rt::wait_req(ri, status);
    return e2 ? e2 : e1;
}
__device__ inline int MPI_Wait(MPI_Request* request, MPI_Status* status) {
    gpumpi::rt::WaveJoin join_;
    if (*request == MPI_REQUEST_NULL) {
        if (status) *status = MPI_Status();
        return MPI_SUCCESS;
    }
    const int idx = gpumpi::rt::req_index(*request);
    if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
    *request = MPI_REQUEST_NULL;
    return gpumpi::rt::wait_req(idx, status);
}
__device__ inline int MPI_Test(MPI_Request* request, int* flag, MPI_Status* status) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    if (*request == MPI_REQUEST_NULL) {
        *flag = 1;
        if (status) *status = MPI_Status();
        return MPI_SUCCESS;
    }
    const int idx = req_index(*request);
    if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
    progress();
    if (reqs()[idx].state == RS_COMPLETE) {
        *flag = 1;
        *request = MPI_REQUEST_NULL;
        fill_status(idx, status);
        const int err = reqs()[idx].st_error;
        req_release(idx);
        return err;
    }
    *flag = 0;
    // A test-polling loop must end once the world is aborted.
    if (++ctx()->test_polls >= ctx()->spin_period) {
        ctx()->test_polls = 0;
        if (aborted()) {
            *flag = 1;
            return MPI_ERR_OTHER;
        }
    }
    return MPI_SUCCESS;
}
__device__ inline int MPI_Waitall(int count, MPI_Request reqv[], MPI_Status* statuses) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    Spinner sp;
    for (;;) {
        bool all = true;
        for (int i = 0; i < count; i++) {
            if (reqv[i] == MPI_REQUEST_NULL) continue;
            const int idx = req_index(reqv[i]);
            if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
            if (reqs()[idx].state != RS_COMPLETE) {
                all = false;
                break;
            }
        }
        if (all) break;
        progress();
        if (!sp.keep_going()) return MPI_ERR_OTHER;
    }
    int err = MPI_SUCCESS;
    for (int i = 0; i < count; i++) {
        MPI_Status* st = statuses ? &statuses[i] : nullptr;
        if (reqv[i] == MPI_REQUEST_NULL) {
            if (st) *st = MPI_Status();
            continue;
        }
        const int idx = req_index(reqv[i]);
        fill_status(idx, st);
        if (reqs()[idx].st_error) err = reqs()[idx].st_error;
        req_release(idx);
        reqv[i] = MPI_REQUEST_NULL;
    }
    return err;
}
__device__ inline int MPI_Testall(int count, MPI_Request reqv[], int* flag, MPI_Status statuses[]) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    progress();
    for (int i = 0; i < count; i++) {
        if (reqv[i] == MPI_REQUEST_NULL) continue;
        const int idx = req_index(reqv[i]);
        if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
        if (reqs()[idx].state != RS_COMPLETE) {
            *flag = 0;
            return MPI_SUCCESS;
        }
    }
    *flag = 1;
    return MPI_Waitall(count, reqv, statuses);
}
__device__ inline int MPI_Waitany(int count, MPI_Request reqv[], int* index, MPI_Status* status) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    Spinner sp;
    for (;;) {
        bool any_active = false;
        for (int i = 0; i < count; i++) {
            if (reqv[i] == MPI_REQUEST_NULL) continue;
            any_active = true;
            const int idx = req_index(reqv[i]);
            if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
            if (reqs()[idx].state == RS_COMPLETE) {
                *index = i;
                fill_status(idx, status);
                const int err = reqs()[idx].st_error;
                req_release(idx);
                reqv[i] = MPI_REQUEST_NULL;
                return err;
            }
        }
        if (!any_active) {
            *index = MPI_UNDEFINED;
            return MPI_SUCCESS;
        }
        progress();
        if (!sp.keep_going()) return MPI_ERR_OTHER;
    }
}
__device__ inline int MPI_Waitsome(int incount, MPI_Request reqv[], int* outcount, int indices[],
                                   MPI_Status statuses[]) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    Spinner sp;
    for (;;) {
        int n = 0;
        bool any_active = false;
        for (int i = 0; i < incount; i++) {
            if (reqv[i] == MPI_REQUEST_NULL) continue;
            any_active = true;
            const int idx = req_index(reqv[i]);
            if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
            if (reqs()[idx].state == RS_COMPLETE) {
                indices[n] = i;
                fill_status(idx, statuses ? &statuses[n] : nullptr);
                req_release(idx);
                reqv[i] = MPI_REQUEST_NULL;
                n++;
            }
        }
        if (!any_active) {
            *outcount = MPI_UNDEFINED;
            return MPI_SUCCESS;
        }
        if (n > 0) {
            *outcount = n;
            return MPI_SUCCESS;
        }
        progress();
        if (!sp.keep_going()) return MPI_ERR_OTHER;
    }
}
// Local and non-blocking: the handle goes away at once, the operation still
// completes, and progress releases the request when it has (mpi.cu:608-617
// freed the request object; here the pool entry is detached).
__device__ inline int MPI_Request_free(MPI_Request* request) {
    gpumpi::rt::WaveJoin join_;
    using namespace gpumpi::rt;
    if (*request == MPI_REQUEST_NULL) return MPI_SUCCESS;
    const int idx = req_index(*request);
    if (idx < 0) return GPUMPI_FAIL(MPI_ERR_REQUEST, "invalid request handle");
    *request = MPI_REQUEST_NULL;
    Req* q = &reqs()[idx];
    if (q->state == RS_COMPLETE) {
        req_release(idx);
        return MPI_SUCCESS;
    }
    q->flags = (uint8_t)(q->flags | RF_DETACHED);
    ctx()->detached++;
    return MPI_SUCCESS;
}

// ---------------------------------------------------------------------------
// Collectives (mpi.cu:86-235, 276-282, 292-541)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Barrier(MPI_Comm comm) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Barrier on MPI_COMM_NULL");
    CommShared* c = comm_get(comm);
    return comm_barrier(c, c->size) ? MPI_SUCCESS : MPI_ERR_OTHER;
}
__device__ inline int MPI_Bcast(void* buffer, int count, MPI_Datatype dt, int root, MPI_Comm comm) {
    return gpumpi::rt::coll_bcast(buffer, count, dt, root, comm);
}
__device__ inline int MPI_Reduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype dt, MPI_Op op,
                                 int root, MPI_Comm comm) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Reduce on MPI_COMM_NULL");
    if (root < 0 || root >= comm_get(comm)->size) return GPUMPI_FAIL(MPI_ERR_RANK, "invalid root in MPI_Reduce");
    return reduce_common(sendbuf, recvbuf, count, dt, op, root, comm);
}
__device__ inline int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype dt, MPI_Op op,
                                    MPI_Comm comm) {
    return gpumpi::rt::reduce_common(sendbuf, recvbuf, count, dt, op, -1, comm);
}
__device__ inline int MPI_Gather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                 int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm) {
    return gpumpi::rt::coll_gather(false, sendbuf, sendcount, sendtype, recvbuf, recvcount, nullptr, nullptr,
                                   recvtype, root, comm);
}
__device__ inline int MPI_Gatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                  const int recvcounts[], const int displs[], MPI_Datatype recvtype, int root,
                                  MPI_Comm comm) {
    return gpumpi::rt::coll_gather(true, sendbuf, sendcount, sendtype, recvbuf, 0, recvcounts, displs, recvtype,
                                   root, comm);
}
__device__ inline int MPI_Scatter(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                  int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm) {
    return gpumpi::rt::coll_scatter(false, sendbuf, sendcount, nullptr, nullptr, sendtype, recvbuf, recvcount,
                                    recvtype, root, comm);
}
__device__ inline int MPI_Scatterv(const void* sendbuf, const int sendcounts[], const int displs[],
                                   MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype,
                                   int root, MPI_Comm comm) {
    return gpumpi::rt::coll_scatter(true, sendbuf, 0, sendcounts, displs, sendtype, recvbuf, recvcount, recvtype,
                                    root, comm);
}
__device__ inline int MPI_Allgather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                    int recvcount, MPI_Datatype recvtype, MPI_Comm comm) {
    return gpumpi::rt::coll_segments(gpumpi::rt::SEG_ALLGATHER, sendbuf, sendcount, nullptr, nullptr, sendtype,
                                     recvbuf, recvcount, nullptr, nullptr, recvtype, comm);
}
__device__ inline int MPI_Allgatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                     const int recvcounts[], const int displs[], MPI_Datatype recvtype,
                                     MPI_Comm comm) {
    return gpumpi::rt::coll_segments(gpumpi::rt::SEG_ALLGATHERV, sendbuf, sendcount, nullptr, nullptr, sendtype,
                                     recvbuf, 0, recvcounts, displs, recvtype, comm);
}
__device__ inline int MPI_Alltoall(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                                   int recvcount, MPI_Datatype recvtype, MPI_Comm comm) {
    return gpumpi::rt::coll_segments(gpumpi::rt::SEG_ALLTOALL, sendbuf, sendcount, nullptr, nullptr, sendtype,
                                     recvbuf, recvcount, nullptr, nullptr, recvtype, comm);
}
__device__ inline int MPI_Alltoallv(const void* sendbuf, const int sendcounts[], const int sdispls[],
                                    MPI_Datatype sendtype, void* recvbuf, const int recvcounts[],
                                    const int rdispls[], MPI_Datatype recvtype, MPI_Comm comm) {
    return gpumpi::rt::coll_segments(gpumpi::rt::SEG_ALLTOALLV, sendbuf, 0, sendcounts, sdispls, sendtype, recvbuf,
                                     0, recvcounts, rdispls, recvtype, comm);
}

// ---------------------------------------------------------------------------
// Communicators (communicator.cu:48-221) and groups (group.cu:53-122)
// ---------------------------------------------------------------------------
__device__ inline int MPI_Comm_size(MPI_Comm comm, int* size) {
    if (!gpumpi::rt::comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Comm_size on MPI_COMM_NULL");
    *size = gpumpi::rt::comm_get(comm)->size;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Comm_rank(MPI_Comm comm, int* rank) {
    if (!gpumpi::rt::comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Comm_rank on MPI_COMM_NULL");
    *rank = gpumpi::rt::comm_rank(gpumpi::rt::comm_get(comm));
    return MPI_SUCCESS;
}
__device__ inline int MPI_Comm_split(MPI_Comm comm, int color, int key, MPI_Comm* newcomm) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Comm_split on MPI_COMM_NULL");
    CommShared* c = comm_get(comm);
    if (color < 0) color = MPI_UNDEFINED;
    *newcomm = comm_split_impl(c, comm_rank(c), color, key);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Comm_dup(MPI_Comm comm, MPI_Comm* newcomm) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return MPI_FAILURE;  // communicator.cu:98-100
    CommShared* c = comm_get(comm);
    const int r = comm_rank(c);
    *newcomm = comm_split_impl(c, r, 0, r);
    return MPI_SUCCESS;
}
// MPI_Comm_free (communicator.cu:80-94 frees on the last reference; freeing
// MPI_COMM_WORLD or MPI_COMM_NULL only nulls the handle).  Collective over the
// members: each completes its own requests on the communicator first (their
// status sources are translated now, while the rank maps exist), then
// counts itself out on the communicator's barrier word; the last member out
// returns the block to the pool and the context id to the free ids.
__device__ inline int MPI_Comm_free(MPI_Comm* comm) {
    using namespace gpumpi::rt;
    WaveJoin join_;
    const MPI_Comm h = *comm;
    *comm = MPI_COMM_NULL;
    if (h == MPI_COMM_NULL || h == MPI_COMM_WORLD || h == MPI_COMM_SELF) return MPI_SUCCESS;
    CommShared* c = comm_get(h);
    if (c->kind != 2) return MPI_SUCCESS;
    RankLds* L = ctx();
    Spinner sp;
    for (;;) {
        bool pending = false;
        for (int idx = 0; idx < L->R; idx++) {
            Req* q = &reqs()[idx];
            if (q->kind == REQ_FREE || (MPI_Comm)(uintptr_t)q->comm != h) continue;
            if (q->state != RS_COMPLETE) {
                pending = true;
                continue;
            }
            if (q->kind == REQ_RECV && q->st_source >= 0) q->st_source = comm_rank_of_world(c, q->st_source);
            q->comm = (uint64_t)(uintptr_t)MPI_COMM_WORLD;  // nothing left to translate
        }
        wave_sync();
        if (!pending) break;
        progress();
        if (!sp.keep_going()) return MPI_ERR_OTHER;
    }
    const int size = c->size, id = c->ctx;
    const uint64_t bytes = c->pool_bytes;
    fence_release_agent();
    uint64_t old = 0;
    if (elected()) old = __hip_atomic_fetch_add(&c->bar, (uint64_t)1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    old = uniform_u64(old);
    if ((old + 1) % (uint64_t)size == 0) {  // the last member out
        fence_acquire_agent();
        comm_pool_free((uint8_t*)c, bytes);
        ctx_free(id);
    }
    return MPI_SUCCESS;
}

namespace gpumpi {
namespace rt {
__device__ inline MPI_Group group_alloc(int n) {
    MPI_Group_impl* g = (MPI_Group_impl*)rank_malloc(sizeof(int) * (uint64_t)(n + 2));
    if (g) g->size = n;
    return g;
}
__device__ inline int group_size(MPI_Group g) {
    if (g == MPI_GROUP_WORLD) return ctx()->nranks;
    if (g == MPI_GROUP_EMPTY || g == MPI_GROUP_NULL) return 0;
    return g->size;
}
__device__ inline int group_world(MPI_Group g, int i) {
    if (g == MPI_GROUP_WORLD) return i;
    return g->ranks[i];
}
}  // namespace rt
}  // namespace gpumpi

__device__ inline int MPI_Comm_group(MPI_Comm comm, MPI_Group* group) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Comm_group on MPI_COMM_NULL");
    if (comm == MPI_COMM_WORLD) {
        *group = MPI_GROUP_WORLD;
        return MPI_SUCCESS;
    }
    CommShared* c = comm_get(comm);
    MPI_Group g = group_alloc(c->size);
    if (!g) return MPI_ERR_NO_MEM;
    for (int i = lane_id(); i < c->size; i += 64) g->ranks[i] = comm_world_of(c, i);
    *group = g;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Group_size(MPI_Group group, int* size) {
    *size = gpumpi::rt::group_size(group);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Group_rank(MPI_Group group, int* rank) {
    using namespace gpumpi::rt;
    const int n = group_size(group), me = my_rank();
    int found = 0x7fffffff;
    for (int i = lane_id(); i < n; i += 64)
        if (group_world(group, i) == me && i < found) found = i;
    for (int off = 32; off > 0; off >>= 1) found = min(found, __shfl_xor(found, off));
    found = uniform_i32(found);
    *rank = found == 0x7fffffff ? MPI_UNDEFINED : found;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Group_incl(MPI_Group group, int n, const int ranks[], MPI_Group* newgroup) {
    using namespace gpumpi::rt;
    MPI_Group g = group_alloc(n);
    if (!g) return MPI_ERR_NO_MEM;
    for (int i = 0; i < n; i++) {
        const int v = group_world(group, ranks[i]);
        g->ranks[i] = v;
    }
    *newgroup = g;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Group_range_incl(MPI_Group group, int n, int ranges[][3], MPI_Group* newgroup) {
    using namespace gpumpi::rt;
    int total = 0;
    for (int i = 0; i < n; i++) {
        const int first = ranges[i][0], last = ranges[i][1], stride = ranges[i][2];
        if (stride == 0) return GPUMPI_FAIL(MPI_ERR_ARG, "zero stride in MPI_Group_range_incl");
        if ((stride > 0 && first > last) || (stride < 0 && first < last)) continue;
        total += (last - first) / stride + 1;
    }
    MPI_Group g = group_alloc(total);
    if (!g) return MPI_ERR_NO_MEM;
    int k = 0;
    for (int i = 0; i < n; i++) {
        const int first = ranges[i][0], last = ranges[i][1], stride = ranges[i][2];
        if ((stride > 0 && first > last) || (stride < 0 && first < last)) continue;
        for (int r = first; stride > 0 ? r <= last : r >= last; r += stride) {
            const int v = group_world(group, r);
            g->ranks[k] = v;
            k++;
        }
    }
    *newgroup = g;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Group_free(MPI_Group* group) {
    const MPI_Group g = *group;
    if (g != MPI_GROUP_NULL && g != MPI_GROUP_WORLD && g != MPI_GROUP_EMPTY)
        gpumpi::rt::rank_free(g, sizeof(int) * (uint64_t)(g->size + 2));
    *group = MPI_GROUP_NULL;
    return MPI_SUCCESS;
}
// Comm_create: members of `group` form a new communicator ordered as in the
// group (communicator.cu:74-87), via the split machinery.
__device__ inline int MPI_Comm_create(MPI_Comm comm, MPI_Group group, MPI_Comm* newcomm) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return GPUMPI_FAIL(MPI_ERR_COMM, "MPI_Comm_create on MPI_COMM_NULL");
    CommShared* c = comm_get(comm);
    int grank;
    MPI_Group_rank(group, &grank);
    *newcomm = comm_split_impl(c, comm_rank(c), grank == MPI_UNDEFINED ? MPI_UNDEFINED : 0, grank);
    return MPI_SUCCESS;
}
__device__ inline int MPI_Cart_create(MPI_Comm, int, const int[], const int[], int, MPI_Comm*) {
    return GPUMPI_FAIL(MPI_ERR_COMM, "Cartesian topologies are not supported");
}
__device__ inline int MPI_Cart_sub(MPI_Comm, const int[], MPI_Comm*) {
    return GPUMPI_FAIL(MPI_ERR_COMM, "Cartesian topologies are not supported");
}
// MPI_Dims_create: balanced factorisation of nnodes over the free (zero)
// dimensions, largest factors first.
__device__ inline int MPI_Dims_create(int nnodes, int ndims, int dims[]) {
    int fixed = 1, nfree = 0;
    for (int i = 0; i < ndims; i++) {
        if (dims[i] > 0) fixed *= dims[i];
        else nfree++;
    }
    if (fixed <= 0 || nnodes % fixed) return GPUMPI_FAIL(MPI_ERR_ARG, "MPI_Dims_create: inconsistent dims");
    int rest = nnodes / fixed;
    int f[64];
    int nf = 0;
    for (int p = 2; p * p <= rest; p++)
        while (rest % p == 0 && nf < 64) {
            f[nf++] = p;
            rest /= p;
        }
    if (rest > 1 && nf < 64) f[nf++] = rest;
    int out[16];
    for (int i = 0; i < 16; i++) out[i] = 1;
    const int nd = nfree < 16 ? nfree : 16;
    for (int k = nf - 1; k >= 0 && nd > 0; k--) {
        int best = 0;
        for (int j = 1; j < nd; j++)
            if (out[j] < out[best]) best = j;
        out[best] *= f[k];
    }
    for (int i = 0; i < nd; i++)
        for (int j = i + 1; j < nd; j++)
            if (out[j] > out[i]) {
                int t = out[i];
                out[i] = out[j];
                out[j] = t;
            }
    int k = 0;
    for (int i = 0; i < ndims; i++)
        if (dims[i] <= 0) dims[i] = out[k++];
    return MPI_SUCCESS;
}
// communicator.cu:130-136 (attributes are accepted and ignored there too).
__device__ inline int MPI_Attr_get(MPI_Comm, int, void*, int* flag) {
    if (flag) *flag = 0;
    return MPI_SUCCESS;
}
__device__ inline int MPI_Attr_put(MPI_Comm, int, void*) { return MPI_SUCCESS; }
__device__ inline int MPI_Keyval_create(MPI_Copy_function*, MPI_Delete_function*, int* keyval, void*) {
    *keyval = 0;
    return MPI_SUCCESS;
}
__device__ inline int MPI_NULL_COPY_FN(MPI_Comm, int, void*, void*, void*, int* flag) {
    if (flag) *flag = 0;
    return MPI_SUCCESS;
}
__device__ inline int MPI_NULL_DELETE_FN(MPI_Comm, int, void*, void*) { return MPI_SUCCESS; }

#include "wire.hpp"
#include "io.hpp"
#include "kernel.hpp"
