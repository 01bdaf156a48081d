/*
 * gpumpi -- device-callable MPI for MI355X (gfx950).
 *
 * Drop-in for the device MPI surface of lilyyangyn/dphpc-gpu_mpi:
 *   gpu_libs/gpu_mpi/mpi.cuh:4-137          (p2p, collectives, requests)
 *   gpu_libs/gpu_mpi/mpi_common.cuh:4-9      (return codes)
 *   gpu_libs/gpu_mpi/datatypes.cuh:12-39     (MPI_Datatype enum)
 *   gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33 (datatype order and sizes)
 *   gpu_libs/gpu_mpi/operators.cuh:7-26      (MPI_Op, MPI_User_function)
 *   gpu_libs/gpu_mpi/mpi_operators_list.cuh:4-16 (operator order)
 *   gpu_libs/gpu_mpi/communicator.cuh:6-49   (MPI_Comm and its calls)
 *   gpu_libs/gpu_mpi/group.cuh:4-31          (MPI_Group and its calls)
 *   gpu_libs/gpu_mpi/io.cuh:13-153           (MPI_File_*: amode, error and seek
 *                                             constants, MPI_File by value)
 *
 * Execution model: every MPI rank is one workgroup of a single persistent
 * kernel.  Wave 0 of the workgroup runs the rank's program (all 64 lanes
 * execute the same scalar code); the other waves of the workgroup join the
 * data movement of MPI calls.  All functions below are called from the
 * rank's program, i.e. from wave 0, with wave-uniform arguments.
 *
 * Differences from the reference that a program can observe (all follow the
 * MPI standard where the reference did not):
 *   - MPI_Sendrecv moves data (the reference's is a stub, mpi.cu:255-260);
 *   - MPI_Status is filled (source, tag, error, byte count);
 *   - MPI_ANY_SOURCE / MPI_ANY_TAG match messages already queued;
 *   - MPI_Barrier works on every communicator;
 *   - MPI_Wait(MPI_REQUEST_NULL) and MPI_Waitall(MPI_STATUSES_IGNORE) are legal;
 *   - MPI_LXOR is the logical exclusive or (the reference used a != b);
 *   - Reductions of every datatype are a deterministic rank-order fold
 *     (the reference's WORLD double sum used atomicAdd, mpi.cu:148-179);
 *   - MPI-IO: MPI_Offset is 64-bit; offsets and positions count etypes of the
 *     view, as MPI specifies (the reference counted bytes); MPI_File_open
 *     does not truncate an existing file and MPI_MODE_EXCL fails only when
 *     the file exists; MPI_File_delete is local; read_at/write_at return the
 *     element count like read/write;
 *   - MPI_Type_contiguous works (datatypes.cu:10-15 left it unimplemented):
 *     a contiguous type of a basic type, usable wherever a datatype is (in
 *     reductions with user-defined operators, as MPI defines the predefined
 *     ones on basic datatypes only); MPI_Type_vector is still refused;
 *   - MPI_Comm_free is collective over the communicator's members: each
 *     completes its own requests on it first, and the last member out returns
 *     its memory and context id (the reference deleted on the last local
 *     reference); freeing MPI_COMM_WORLD, MPI_COMM_SELF or MPI_COMM_NULL only
 *     nulls the handle, as in the reference.
 */
#ifndef GPUMPI_MPI_H
#define GPUMPI_MPI_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

/* ---- constants: mpi.cuh:4-9, mpi_common.cuh:4-9, mpi.cuh:130-137 ----- */
#define MPI_MAX_PROCESSOR_NAME 256
#define MPI_ANY_TAG (-1)
#define MPI_ANY_SOURCE (-1)
#define MPI_PROC_NULL (-3)
#define MPI_TAG_UB ((1 << 15) - 1)

#define MPI_SUCCESS 0
#define MPI_FAILURE -1
#define MPI_ERR_OTHER 1
#define MPI_UNDEFINED -1
/* Extra error classes (values above the reference's IO codes, io.cuh:28-44). */
#define MPI_ERR_TRUNCATE 20
#define MPI_ERR_COUNT 21
#define MPI_ERR_TYPE 22
#define MPI_ERR_OP 23
#define MPI_ERR_RANK 24
#define MPI_ERR_COMM 25
#define MPI_ERR_ARG 26
#define MPI_ERR_NO_MEM 27
#define MPI_ERR_REQUEST 28
#define MPI_ERR_TAG 29
#define MPI_ERR_BUFFER 30
#define MPI_ERR_INTERN 31

/* MPI-IO error classes, io.cuh:28-44 (their values collide with
 * MPI_ERR_OTHER and each other's neighbours there too). */
#define MPI_ERR_FILE 1
#define MPI_ERR_NOT_SAME 2
#define MPI_ERR_AMODE 3
#define MPI_ERR_UNSUPPORTED_DATAREP 4
#define MPI_ERR_UNSUPPORTED_OPERATION 5
#define MPI_ERR_NO_SUCH_FILE 6
#define MPI_ERR_FILE_EXISTS 7
#define MPI_ERR_BAD_FILE 8
#define MPI_ERR_ACCESS 9
#define MPI_ERR_NO_SPACE 10
#define MPI_ERR_QUOTA 11
#define MPI_ERR_READ_ONLY 12
#define MPI_ERR_FILE_IN_USE 13
#define MPI_ERR_DUP_DATAREP 14
#define MPI_ERR_CONVERSION 15
#define MPI_ERR_IO 16

/* amode bits, io.cuh:13-21 */
#define MPI_MODE_RDONLY 1
#define MPI_MODE_RDWR 2
#define MPI_MODE_WRONLY 4
#define MPI_MODE_CREATE 8
#define MPI_MODE_EXCL 16
#define MPI_MODE_DELETE_ON_CLOSE 32
#define MPI_MODE_UNIQUE_OPEN 64
#define MPI_MODE_SEQUENTIAL 128
#define MPI_MODE_APPEND 256
/* io.cuh:50-52, 76 */
#define MPI_SEEK_SET 1
#define MPI_SEEK_CUR 2
#define MPI_SEEK_END 3
#define MPI_DISPLACEMENT_CURRENT (-1)

#define MPI_THREAD_SINGLE 1
#define MPI_THREAD_FUNNELED 2
#define MPI_THREAD_SERIALIZED 3
#define MPI_THREAD_MULTIPLE 4

/* ---- datatypes: mpi_types_list.cuh:3-33, same order ------------------ */
enum MPI_Datatype {
    MPI_CHAR, MPI_SHORT, MPI_INT, MPI_LONG, MPI_LONG_LONG_INT, MPI_LONG_LONG,
    MPI_SIGNED_CHAR, MPI_UNSIGNED_CHAR, MPI_UNSIGNED_SHORT, MPI_UNSIGNED,
    MPI_UNSIGNED_LONG, MPI_UNSIGNED_LONG_LONG, MPI_FLOAT, MPI_DOUBLE,
    MPI_LONG_DOUBLE, MPI_WCHAR, MPI_C_BOOL, MPI_INT8_T, MPI_INT16_T,
    MPI_INT32_T, MPI_INT64_T, MPI_UINT8_T, MPI_UINT16_T, MPI_UINT32_T,
    MPI_UINT64_T, MPI_C_COMPLEX, MPI_C_FLOAT_COMPLEX, MPI_C_DOUBLE_COMPLEX,
    MPI_C_LONG_DOUBLE_COMPLEX, MPI_BYTE,
    GPUMPI_NUM_BASIC_TYPES,
    MPI_DATATYPE_NULL = -1,
    /* MPI_Type_contiguous handles: GPUMPI_CONTIGUOUS_FLAG | count << 5 | basic
     * type (the same value on every rank, no table behind it) */
    GPUMPI_CONTIGUOUS_FLAG = 1 << 30,
    GPUMPI_DATATYPE_MAX = 0x7fffffff
};

/* ---- handles --------------------------------------------------------- */
struct MPI_Comm_impl;
typedef struct MPI_Comm_impl* MPI_Comm;      /* communicator.cuh:6-8 */
struct MPI_Group_impl;
typedef struct MPI_Group_impl* MPI_Group;    /* group.cuh:4-5 */
struct MPI_Op_impl;
typedef struct MPI_Op_impl* MPI_Op;          /* operators.cuh:7-8 */
struct MPI_Request_impl;
typedef struct MPI_Request_impl* MPI_Request; /* mpi.cuh:26-29 */

/* Predefined handles are constants, valid before MPI_Init. */
#define MPI_COMM_NULL ((MPI_Comm)0)
#define MPI_COMM_WORLD ((MPI_Comm)(uintptr_t)1)
#define MPI_COMM_SELF ((MPI_Comm)(uintptr_t)2)
#define MPI_GROUP_NULL ((MPI_Group)0)
#define MPI_GROUP_WORLD ((MPI_Group)(uintptr_t)1)
#define MPI_GROUP_EMPTY ((MPI_Group)(uintptr_t)2)
#define MPI_REQUEST_NULL ((MPI_Request)0)
#define MPI_OP_NULL ((MPI_Op)0)

/* Built-in operators in the order of mpi_operators_list.cuh:4-16; the handle
 * value is 1 + the operator's index. */
#define GPUMPI_OP_HANDLE(i) ((MPI_Op)(uintptr_t)(1 + (i)))
#define MPI_MAX GPUMPI_OP_HANDLE(0)
#define MPI_MIN GPUMPI_OP_HANDLE(1)
#define MPI_SUM GPUMPI_OP_HANDLE(2)
#define MPI_PROD GPUMPI_OP_HANDLE(3)
#define MPI_LAND GPUMPI_OP_HANDLE(4)
#define MPI_BAND GPUMPI_OP_HANDLE(5)
#define MPI_LOR GPUMPI_OP_HANDLE(6)
#define MPI_BOR GPUMPI_OP_HANDLE(7)
#define MPI_LXOR GPUMPI_OP_HANDLE(8)
#define MPI_BXOR GPUMPI_OP_HANDLE(9)
#define MPI_MAXLOC GPUMPI_OP_HANDLE(10)
#define MPI_MINLOC GPUMPI_OP_HANDLE(11)
#define GPUMPI_NUM_BUILTIN_OPS 12

/* MPI_IN_PLACE for the send buffer of Reduce/Allreduce/Gather/Allgather. */
#define MPI_IN_PLACE ((void*)(uintptr_t)1)

/* mpi.cuh:19-23 plus the received byte count (MPI_Get_count). */
struct MPI_Status {
    int MPI_SOURCE = MPI_ANY_SOURCE;
    int MPI_TAG = MPI_ANY_TAG;
    int MPI_ERROR = MPI_SUCCESS;
    int gpumpi_count_lo = 0; /* received bytes (low, high 32 bits) */
    int gpumpi_count_hi = 0;
};
#define MPI_STATUS_IGNORE ((MPI_Status*)nullptr)
#define MPI_STATUSES_IGNORE ((MPI_Status*)nullptr)

/* operators.cuh:20 */
typedef void MPI_User_function(void* invec, void* inoutvec, int* len, MPI_Datatype* datatype);

/* mpi.cuh:83-93 */
typedef int MPI_Copy_function(MPI_Comm oldcomm, int keyval, void* extra_state,
                              void* attribute_val_in, void* attribute_val_out, int* flag);
typedef int MPI_Delete_function(MPI_Comm comm, int keyval, void* attribute_val, void* extra_state);

/* ---- the program entry points a converted program defines --------------
 * converter.cpp:52-64 / gpu_main/main.cu:99-100. */
__device__ int __gpu_main(int argc, char* argv[]);

/* ---- environment: mpi.cu:44-75, 142-146, 271-274 ---------------------- */
__device__ int MPI_Init(int* argc, char*** argv);
__device__ int MPI_Init_thread(int* argc, char*** argv, int required, int* provided);
__device__ int MPI_Initialized(int* flag);
__device__ int MPI_Finalize(void);
__device__ int MPI_Abort(MPI_Comm comm, int errorcode);
__device__ int MPI_Get_processor_name(char* name, int* resultlen);
__device__ double MPI_Wtime(void);
__device__ double MPI_Wtick(void);

/* ---- datatypes: mpi.cu:237-245, 283-290 -------------------------------- */
__device__ int MPI_Type_size(MPI_Datatype datatype, int* size);
__device__ int MPI_Type_contiguous(int count, MPI_Datatype oldtype, MPI_Datatype* newtype);
__device__ int MPI_Type_commit(MPI_Datatype* datatype);
__device__ int MPI_Type_free(MPI_Datatype* datatype);
__device__ int MPI_Type_vector(int count, int blocklength, int stride, MPI_Datatype oldtype, MPI_Datatype* newtype);
__device__ int MPI_Get_count(const MPI_Status* status, MPI_Datatype datatype, int* count);

/* ---- operators: operators.cu:241-254 ----------------------------------- */
__device__ int MPI_Op_create(MPI_User_function* user_fn, int commute, MPI_Op* op);
__device__ int MPI_Op_free(MPI_Op* op);

/* ---- point to point: mpi.cu:247-269, 543-617 ---------------------------- */
__device__ int MPI_Send(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                        MPI_Comm comm);
__device__ int MPI_Ssend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                         MPI_Comm comm);
__device__ int MPI_Recv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                        MPI_Comm comm, MPI_Status* status);
__device__ int MPI_Sendrecv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, int dest,
                            int sendtag, void* recvbuf, int recvcount, MPI_Datatype recvtype,
                            int source, int recvtag, MPI_Comm comm, MPI_Status* status);
__device__ int MPI_Isend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                         MPI_Comm comm, MPI_Request* request);
__device__ int MPI_Issend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                          MPI_Comm comm, MPI_Request* request);
__device__ int MPI_Irecv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                         MPI_Comm comm, MPI_Request* request);
__device__ int MPI_Wait(MPI_Request* request, MPI_Status* status);
__device__ int MPI_Test(MPI_Request* request, int* flag, MPI_Status* status);
__device__ int MPI_Waitall(int count, MPI_Request array_of_requests[], MPI_Status* array_of_statuses);
__device__ int MPI_Testall(int count, MPI_Request array_of_requests[], int* flag,
                           MPI_Status array_of_statuses[]);
__device__ int MPI_Waitany(int count, MPI_Request array_of_requests[], int* index, MPI_Status* status);
__device__ int MPI_Waitsome(int incount, MPI_Request array_of_requests[], int* outcount,
                            int array_of_indices[], MPI_Status array_of_statuses[]);
__device__ int MPI_Request_free(MPI_Request* request);

/* ---- collectives: mpi.cu:86-235, 276-282, 292-541 ------------------------ */
__device__ int MPI_Barrier(MPI_Comm comm);
__device__ int MPI_Bcast(void* buffer, int count, MPI_Datatype datatype, int root, MPI_Comm comm);
__device__ int MPI_Reduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype,
                          MPI_Op op, int root, MPI_Comm comm);
__device__ int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype,
                             MPI_Op op, MPI_Comm comm);
__device__ int MPI_Gather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                          int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm);
__device__ int MPI_Gatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                           const int recvcounts[], const int displs[], MPI_Datatype recvtype,
                           int root, MPI_Comm comm);
__device__ int MPI_Scatter(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                           int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm);
__device__ int MPI_Scatterv(const void* sendbuf, const int sendcounts[], const int displs[],
                            MPI_Datatype sendtype, void* recvbuf, int recvcount,
                            MPI_Datatype recvtype, int root, MPI_Comm comm);
__device__ int MPI_Allgather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                             int recvcount, MPI_Datatype recvtype, MPI_Comm comm);
__device__ int MPI_Allgatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                              void* recvbuf, const int recvcounts[], const int displs[],
                              MPI_Datatype recvtype, MPI_Comm comm);
__device__ int MPI_Alltoall(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf,
                            int recvcount, MPI_Datatype recvtype, MPI_Comm comm);
__device__ int MPI_Alltoallv(const void* sendbuf, const int sendcounts[], const int sdispls[],
                             MPI_Datatype sendtype, void* recvbuf, const int recvcounts[],
                             const int rdispls[], MPI_Datatype recvtype, MPI_Comm comm);

/* ---- communicators: communicator.cuh:23-49 ----------------------------- */
__device__ int MPI_Comm_size(MPI_Comm comm, int* size);
__device__ int MPI_Comm_rank(MPI_Comm comm, int* rank);
__device__ int MPI_Comm_group(MPI_Comm comm, MPI_Group* group);
__device__ int MPI_Comm_create(MPI_Comm comm, MPI_Group group, MPI_Comm* newcomm);
__device__ int MPI_Comm_dup(MPI_Comm comm, MPI_Comm* newcomm);
__device__ int MPI_Comm_split(MPI_Comm comm, int color, int key, MPI_Comm* newcomm);
__device__ int MPI_Comm_free(MPI_Comm* comm);
__device__ int MPI_Cart_create(MPI_Comm comm_old, int ndims, const int dims[], const int periods[],
                               int reorder, MPI_Comm* comm_cart);
__device__ int MPI_Cart_sub(MPI_Comm comm, const int remain_dims[], MPI_Comm* comm_new);
__device__ int MPI_Dims_create(int nnodes, int ndims, int dims[]);
__device__ int MPI_Attr_get(MPI_Comm comm, int keyval, void* attribute_val, int* flag);
__device__ int MPI_Attr_put(MPI_Comm comm, int keyval, void* attribute_val);
__device__ int MPI_Keyval_create(MPI_Copy_function* copy_fn, MPI_Delete_function* delete_fn,
                                 int* keyval, void* extra_state);
__device__ int MPI_NULL_COPY_FN(MPI_Comm oldcomm, int keyval, void* extra_state,
                                void* attribute_val_in, void* attribute_val_out, int* flag);
__device__ int MPI_NULL_DELETE_FN(MPI_Comm comm, int keyval, void* attribute_val, void* extra_state);

/* ---- MPI-IO: io.cuh:56-153 --------------------------------------------- */
typedef long long MPI_Offset;
typedef struct MPI_Info_s {
    int gpumpi_unused;
} MPI_Info;
#define MPI_INFO_NULL (MPI_Info{0})
struct gpumpi_file_state;
/* By value, as io.cuh:93-101: `file` is the host's file handle (non-null
 * while open), `st` the positions and views shared by the members. */
typedef struct MPI_File_s {
    MPI_Comm comm;
    void* file;
    int amode;
    struct gpumpi_file_state* st;
} MPI_File;
#define MPI_FILE_NULL (MPI_File{MPI_COMM_NULL, nullptr, 0, nullptr})
__device__ inline bool operator==(const MPI_File& a, const MPI_File& b) { return a.file == b.file && a.st == b.st; }
__device__ inline bool operator!=(const MPI_File& a, const MPI_File& b) { return !(a == b); }

__device__ int MPI_File_open(MPI_Comm comm, const char* filename, int amode, MPI_Info info, MPI_File* fh);
__device__ int MPI_File_close(MPI_File* fh);
__device__ int MPI_File_delete(const char* filename, MPI_Info info);
__device__ int MPI_File_get_size(MPI_File fh, MPI_Offset* size);
__device__ int MPI_File_get_amode(MPI_File fh, int* amode);
__device__ int MPI_File_sync(MPI_File fh);
__device__ int MPI_File_set_view(MPI_File fh, MPI_Offset disp, MPI_Datatype etype, MPI_Datatype filetype,
                                 const char* datarep, MPI_Info info);
__device__ int MPI_File_get_view(MPI_File fh, MPI_Offset* disp, MPI_Datatype* etype, MPI_Datatype* filetype,
                                 char* datarep);
__device__ int MPI_File_read(MPI_File fh, void* buf, int count, MPI_Datatype datatype, MPI_Status* status);
__device__ int MPI_File_write(MPI_File fh, const void* buf, int count, MPI_Datatype datatype, MPI_Status* status);
__device__ int MPI_File_read_at(MPI_File fh, MPI_Offset offset, void* buf, int count, MPI_Datatype datatype,
                                MPI_Status* status);
__device__ int MPI_File_write_at(MPI_File fh, MPI_Offset offset, const void* buf, int count,
                                 MPI_Datatype datatype, MPI_Status* status);
__device__ int MPI_File_read_all(MPI_File fh, void* buf, int count, MPI_Datatype datatype, MPI_Status* status);
__device__ int MPI_File_write_all(MPI_File fh, const void* buf, int count, MPI_Datatype datatype,
                                  MPI_Status* status);
__device__ int MPI_File_read_shared(MPI_File fh, void* buf, int count, MPI_Datatype datatype, MPI_Status* status);
__device__ int MPI_File_write_shared(MPI_File fh, const void* buf, int count, MPI_Datatype datatype,
                                     MPI_Status* status);
__device__ int MPI_File_seek(MPI_File fh, MPI_Offset offset, int whence);
__device__ int MPI_File_get_position(MPI_File fh, MPI_Offset* offset);
__device__ int MPI_File_seek_shared(MPI_File fh, MPI_Offset offset, int whence);
__device__ int MPI_File_get_position_shared(MPI_File fh, MPI_Offset* offset);

/* ---- groups: group.cuh:21-31 ------------------------------------------- */
__device__ int MPI_Group_size(MPI_Group group, int* size);
__device__ int MPI_Group_rank(MPI_Group group, int* rank);
__device__ int MPI_Group_incl(MPI_Group group, int n, const int ranks[], MPI_Group* newgroup);
__device__ int MPI_Group_range_incl(MPI_Group group, int n, int ranges[][3], MPI_Group* newgroup);
__device__ int MPI_Group_free(MPI_Group* group);

/* The implementation is header-only so every MPI call inlines into the rank
 * program's kernel. */
#include "gpumpi_rt/mpi_impl.hpp"

#endif /* GPUMPI_MPI_H */
