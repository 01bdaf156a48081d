// MPI-IO on the host service channel (svc.hpp).
//
// Follows the contract of gpu_libs/gpu_mpi/io.cu:137-559 and the behaviour
// its tests observe (tests/test_io_file_open.cu, test_io_file_access.cu,
// test_io_file_write.cu): MPI_File is a value whose `file` is the host's
// handle; amode is checked in the standard's order (io.cu:149-161); reads and
// writes return the element count; a read without read access returns
// MPI_ERR_AMODE, one on a sequential file MPI_ERR_UNSUPPORTED_OPERATION;
// the default view is (0, MPI_CHAR, MPI_CHAR, "native").  Rank 0 of the
// file's communicator opens and closes the host file and owns the shared
// state; every member moves its own data.  Where the reference departs from
// the MPI standard this follows the standard (see include/gpumpi/mpi.h).
// Errors are return values only: an I/O error does not abort the launch.
#pragma once

#include "svc.hpp"

struct gpumpi_file_view {
    MPI_Offset disp;  // bytes
    MPI_Datatype etype;
    MPI_Datatype filetype;
    char datarep[16];
};
// Shared by the members of an open file (rank 0's heap): positions count
// etypes of each member's view.
struct gpumpi_file_state {
    int size;
    int pad;
    unsigned long long shared_pos;  // in etypes of the (common) view
    // then: MPI_Offset pos[size]; gpumpi_file_view view[size]
};

namespace gpumpi {
namespace rt {

__device__ inline MPI_Offset* file_pos(gpumpi_file_state* st) { return (MPI_Offset*)(st + 1); }
__device__ inline gpumpi_file_view* file_view(gpumpi_file_state* st) {
    return (gpumpi_file_view*)(file_pos(st) + st->size);
}
__device__ inline uint64_t file_state_bytes(int size) {
    return sizeof(gpumpi_file_state) + (uint64_t)size * (sizeof(MPI_Offset) + sizeof(gpumpi_file_view));
}

// amode legality, io.cu:149-161 (MPI-3.1 p495).
__device__ inline int file_check_amode(int amode) {
    const int rw = ((amode & MPI_MODE_RDONLY) ? 1 : 0) + ((amode & MPI_MODE_RDWR) ? 1 : 0) +
                   ((amode & MPI_MODE_WRONLY) ? 1 : 0);
    if (rw != 1) return MPI_ERR_AMODE;
    if ((amode & (MPI_MODE_CREATE | MPI_MODE_EXCL)) && (amode & MPI_MODE_RDONLY)) return MPI_ERR_AMODE;
    if ((amode & MPI_MODE_RDWR) && (amode & MPI_MODE_SEQUENTIAL)) return MPI_ERR_AMODE;
    return MPI_SUCCESS;
}

__device__ inline bool file_valid(const MPI_File& fh) { return fh.file != nullptr && fh.st != nullptr; }

// One host transfer of at most kSvcDataBytes at byte offset `off`; returns
// the bytes moved or -1 (with *err set).
__device__ inline int64_t file_xfer(void* host_file, bool write, uint64_t off, uint8_t* buf, uint64_t bytes,
                                    int* err) {
    SvcSlot* s = svc_slot();
    if (!s) {
        *err = MPI_ERR_IO;
        return -1;
    }
    uint64_t done = 0;
    while (done < bytes) {
        const uint64_t n = bytes - done < (uint64_t)kSvcDataBytes ? bytes - done : (uint64_t)kSvcDataBytes;
        s->l[0] = (int64_t)(uintptr_t)host_file;
        s->l[1] = (int64_t)(off + done);
        s->i[0] = (int32_t)n;
        if (write) svc_copy(s->data, buf + done, n);
        if (!svc_call(s, write ? SVC_WRITE : SVC_READ)) {
            *err = MPI_ERR_IO;
            return -1;
        }
        const int64_t got = s->i[1];
        const int e = s->i[2];
        if (!write && got > 0) svc_copy(buf + done, s->data, (uint64_t)got);
        svc_release(s);
        if (e != MPI_SUCCESS) {
            *err = e;
            return -1;
        }
        done += (uint64_t)got;
        if ((uint64_t)got < n) break;  // end of file
    }
    *err = MPI_SUCCESS;
    return (int64_t)done;
}

__device__ inline int64_t file_host_size(void* host_file) {
    SvcSlot* s = svc_slot();
    if (!s) return -1;
    s->l[0] = (int64_t)(uintptr_t)host_file;
    if (!svc_call(s, SVC_SIZE)) return -1;
    const int64_t sz = s->i[1] == MPI_SUCCESS ? s->l[1] : -1;
    svc_release(s);
    return sz;
}

__device__ inline void file_status(MPI_Status* status, uint64_t bytes) {
    if (!status) return;
    status->MPI_SOURCE = MPI_ANY_SOURCE;
    status->MPI_TAG = MPI_ANY_TAG;
    status->MPI_ERROR = MPI_SUCCESS;
    status->gpumpi_count_lo = (int)(uint32_t)bytes;
    status->gpumpi_count_hi = (int)(uint32_t)(bytes >> 32);
}

// Read or write `count` elements at element offset `at` of this member's
// view; returns the elements moved, or an error code through *err.
__device__ inline int file_rw(const MPI_File& fh, bool write, MPI_Offset at, void* buf, int count,
                              MPI_Datatype dt, MPI_Status* status, int* err) {
    const int es = type_size(dt);
    if (es <= 0 || count < 0) {
        *err = es <= 0 ? MPI_ERR_TYPE : MPI_ERR_COUNT;
        return 0;
    }
    const gpumpi_file_view v = file_view(fh.st)[comm_rank(comm_get(fh.comm))];
    const int ees = type_size(v.etype);
    const uint64_t off = (uint64_t)v.disp + (uint64_t)at * (uint64_t)ees;
    const int64_t moved = file_xfer(fh.file, write, off, (uint8_t*)buf, (uint64_t)count * (uint64_t)es, err);
    if (moved < 0) return 0;
    file_status(status, (uint64_t)moved);
    return (int)(moved / es);
}

struct FileOpenReply {
    int err;
    int pad;
    void* file;
    gpumpi_file_state* st;
};

}  // namespace rt
}  // namespace gpumpi

// io.cu:137-234
__device__ inline int MPI_File_open(MPI_Comm comm, const char* filename, int amode, MPI_Info, MPI_File* fh) {
    using namespace gpumpi::rt;
    if (!comm_valid(comm)) return MPI_ERR_COMM;
    int rank = 0, size = 1;
    MPI_Comm_rank(comm, &rank);
    MPI_Comm_size(comm, &size);
    FileOpenReply rep{file_check_amode(amode), 0, nullptr, nullptr};
    if (rank == 0 && rep.err == MPI_SUCCESS) {
        SvcSlot* s = svc_slot();
        if (!s || !filename || svc_put_string(s, 0, filename) < 0) {
            rep.err = !s ? MPI_ERR_IO : MPI_ERR_BAD_FILE;
        } else {
            s->i[0] = amode;
            if (!svc_call(s, SVC_OPEN)) {
                rep.err = MPI_ERR_IO;
            } else {
                rep.err = s->i[1];
                rep.file = (void*)(uintptr_t)s->l[0];
                const int64_t fsize = s->l[1];
                svc_release(s);
                if (rep.err == MPI_SUCCESS) {
                    gpumpi_file_state* st = (gpumpi_file_state*)rank_malloc(file_state_bytes(size));
                    if (!st) {
                        rep.err = MPI_ERR_NO_SPACE;
                    } else {
                        if (elected()) {
                            st->size = size;
                            st->shared_pos = 0;
                        }
                        wave_sync();
                        // MPI_MODE_APPEND: every file pointer starts at the end (p494).
                        const MPI_Offset init = (amode & MPI_MODE_APPEND) ? (MPI_Offset)fsize : 0;
                        for (int r = lane_id(); r < size; r += 64) {
                            file_pos(st)[r] = init;
                            gpumpi_file_view& v = file_view(st)[r];
                            v.disp = 0;
                            v.etype = MPI_CHAR;
                            v.filetype = MPI_CHAR;
                            for (int k = 0; k < 16; k++) v.datarep[k] = "native"[k < 7 ? k : 6];
                        }
                        if (elected()) st->shared_pos = (unsigned long long)init;
                        wave_sync();
                        rep.st = st;
                    }
                }
            }
        }
    }
    const int berr = MPI_Bcast(&rep, (int)sizeof rep, MPI_BYTE, 0, comm);
    if (berr) return berr;
    if (rep.err != MPI_SUCCESS) {
        *fh = MPI_FILE_NULL;
        return rep.err;
    }
    *fh = MPI_File{comm, rep.file, amode, rep.st};
    return MPI_SUCCESS;
}

// io.cu:272-292
__device__ inline int MPI_File_close(MPI_File* fh) {
    using namespace gpumpi::rt;
    if (!fh || !file_valid(*fh)) return MPI_ERR_FILE;
    MPI_Barrier(fh->comm);
    int rank = 0;
    MPI_Comm_rank(fh->comm, &rank);
    int err = MPI_SUCCESS;
    if (rank == 0) {
        SvcSlot* s = svc_slot();
        if (!s) {
            err = MPI_ERR_IO;
        } else {
            s->l[0] = (int64_t)(uintptr_t)fh->file;
            s->i[0] = (fh->amode & MPI_MODE_DELETE_ON_CLOSE) ? 1 : 0;
            if (!svc_call(s, SVC_CLOSE)) {
                err = MPI_ERR_IO;
            } else {
                err = s->i[1];
                svc_release(s);
            }
        }
        rank_free(fh->st, file_state_bytes(fh->st->size));
    }
    MPI_Bcast(&err, 1, MPI_INT, 0, fh->comm);
    *fh = MPI_FILE_NULL;
    return err;
}

// io.cu:238-253 (local here, as the standard has it).
__device__ inline int MPI_File_delete(const char* filename, MPI_Info) {
    using namespace gpumpi::rt;
    SvcSlot* s = svc_slot();
    if (!s) return MPI_ERR_IO;
    if (!filename || svc_put_string(s, 0, filename) < 0) return MPI_ERR_BAD_FILE;
    if (!svc_call(s, SVC_DELETE)) return MPI_ERR_IO;
    const int err = s->i[1];
    svc_release(s);
    return err;
}

// io.cu:255-270: collective, rank 0 asks the host and broadcasts.
__device__ inline int MPI_File_get_size(MPI_File fh, MPI_Offset* size) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    MPI_Barrier(fh.comm);
    int rank = 0;
    MPI_Comm_rank(fh.comm, &rank);
    long long sz = 0;
    if (rank == 0) sz = file_host_size(fh.file);
    MPI_Bcast(&sz, 1, MPI_LONG_LONG, 0, fh.comm);
    if (sz < 0) return MPI_ERR_IO;
    *size = sz;
    return MPI_SUCCESS;
}

__device__ inline int MPI_File_get_amode(MPI_File fh, int* amode) {
    if (!gpumpi::rt::file_valid(fh)) return MPI_ERR_FILE;
    *amode = fh.amode;
    return MPI_SUCCESS;
}

__device__ inline int MPI_File_sync(MPI_File fh) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    SvcSlot* s = svc_slot();
    if (!s) return MPI_ERR_IO;
    s->l[0] = (int64_t)(uintptr_t)fh.file;
    if (!svc_call(s, SVC_SYNC)) return MPI_ERR_IO;
    const int err = s->i[1];
    svc_release(s);
    return err;
}

// io.cu:297-314: the member's own view; the individual pointer and (rank 0)
// the shared pointer restart at the view's beginning.
__device__ inline int MPI_File_set_view(MPI_File fh, MPI_Offset disp, MPI_Datatype etype, MPI_Datatype filetype,
                                        const char* datarep, MPI_Info) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    // a dense view only: the filetype is etype or a contiguous run of its elements
    if (type_size(etype) <= 0 || type_size(filetype) <= 0 || type_base(filetype) != type_base(etype) ||
        type_size(filetype) % type_size(etype) != 0)
        return MPI_ERR_TYPE;
    const char* native = "native";
    if (datarep) {
        int k = 0;
        while (native[k] && datarep[k] == native[k]) k++;
        if (native[k] != 0 || datarep[k] != 0) return MPI_ERR_UNSUPPORTED_DATAREP;
    }
    const int me = comm_rank(comm_get(fh.comm));
    if (disp == MPI_DISPLACEMENT_CURRENT) {
        if (!(fh.amode & MPI_MODE_SEQUENTIAL)) return MPI_ERR_ARG;
        const gpumpi_file_view& cur = file_view(fh.st)[me];
        disp = cur.disp + (MPI_Offset)fh.st->shared_pos * type_size(cur.etype);
    }
    if (disp < 0) return MPI_ERR_ARG;
    MPI_Barrier(fh.comm);  // no member still reads the old views or pointers
    if (elected()) {
        gpumpi_file_view& v = file_view(fh.st)[me];
        v.disp = disp;
        v.etype = etype;
        v.filetype = filetype;
        for (int k = 0; k < 16; k++) v.datarep[k] = native[k < 7 ? k : 6];
        file_pos(fh.st)[me] = 0;
        if (me == 0) fh.st->shared_pos = 0;
    }
    wave_sync();
    MPI_Barrier(fh.comm);
    return MPI_SUCCESS;
}

// io.cu:316-331
__device__ inline int MPI_File_get_view(MPI_File fh, MPI_Offset* disp, MPI_Datatype* etype, MPI_Datatype* filetype,
                                        char* datarep) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    const gpumpi_file_view v = file_view(fh.st)[comm_rank(comm_get(fh.comm))];
    *disp = v.disp;
    *etype = v.etype;
    *filetype = v.filetype;
    if (datarep)
        for (int k = 0; k < 16; k++) {
            datarep[k] = v.datarep[k];
            if (v.datarep[k] == 0) break;
        }
    return MPI_SUCCESS;
}

// io.cu:336-371: offsets count etypes of the member's view.
__device__ inline int MPI_File_seek(MPI_File fh, MPI_Offset offset, int whence) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (fh.amode & MPI_MODE_SEQUENTIAL) return MPI_ERR_UNSUPPORTED_OPERATION;
    const int me = comm_rank(comm_get(fh.comm));
    MPI_Offset* pos = &file_pos(fh.st)[me];
    MPI_Offset np;
    if (whence == MPI_SEEK_SET) {
        np = offset;
    } else if (whence == MPI_SEEK_CUR) {
        np = *pos + offset;
    } else if (whence == MPI_SEEK_END) {
        const gpumpi_file_view v = file_view(fh.st)[me];
        const int64_t sz = file_host_size(fh.file);
        if (sz < 0) return MPI_ERR_IO;
        const int64_t in_view = sz - v.disp > 0 ? sz - v.disp : 0;
        np = (in_view + type_size(v.etype) - 1) / type_size(v.etype) + offset;
    } else {
        return MPI_ERR_ARG;
    }
    if (np < 0) return MPI_ERR_UNSUPPORTED_OPERATION;  // io.cu:347 (p521 l11)
    if (elected()) *pos = np;
    wave_sync();
    return MPI_SUCCESS;
}

// io.cu:373-378
__device__ inline int MPI_File_get_position(MPI_File fh, MPI_Offset* offset) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    *offset = file_pos(fh.st)[comm_rank(comm_get(fh.comm))];
    return MPI_SUCCESS;
}

// io.cu:380-418: reads at the individual pointer and advances it.
__device__ inline int MPI_File_read(MPI_File fh, void* buf, int count, MPI_Datatype datatype, MPI_Status* status) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (!(fh.amode & MPI_MODE_RDONLY) && !(fh.amode & MPI_MODE_RDWR)) return MPI_ERR_AMODE;
    if (fh.amode & MPI_MODE_SEQUENTIAL) return MPI_ERR_UNSUPPORTED_OPERATION;  // p514 l43
    const int me = comm_rank(comm_get(fh.comm));
    MPI_Offset* pos = &file_pos(fh.st)[me];
    int err = MPI_SUCCESS;
    const int n = file_rw(fh, false, *pos, buf, count, datatype, status, &err);
    if (err) return err;
    const int ees = type_size(file_view(fh.st)[me].etype);
    if (elected()) *pos += (MPI_Offset)n * type_size(datatype) / ees;
    wave_sync();
    return n;
}

// io.cu:421-455: writes at the individual pointer and advances it.
__device__ inline int MPI_File_write(MPI_File fh, const void* buf, int count, MPI_Datatype datatype,
                                     MPI_Status* status) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (fh.amode & MPI_MODE_RDONLY) return MPI_ERR_READ_ONLY;
    const int me = comm_rank(comm_get(fh.comm));
    MPI_Offset* pos = &file_pos(fh.st)[me];
    int err = MPI_SUCCESS;
    const int n = file_rw(fh, true, *pos, (void*)buf, count, datatype, status, &err);
    if (err) return err;
    const int ees = type_size(file_view(fh.st)[me].etype);
    if (elected()) *pos += (MPI_Offset)n * type_size(datatype) / ees;
    wave_sync();
    return n;
}

// io.cu:457-505: explicit offsets; the individual pointer does not move.
__device__ inline int MPI_File_read_at(MPI_File fh, MPI_Offset offset, void* buf, int count, MPI_Datatype datatype,
                                       MPI_Status* status) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (!(fh.amode & MPI_MODE_RDONLY) && !(fh.amode & MPI_MODE_RDWR)) return MPI_ERR_AMODE;
    if (fh.amode & MPI_MODE_SEQUENTIAL) return MPI_ERR_UNSUPPORTED_OPERATION;
    if (offset < 0) return MPI_ERR_ARG;
    int err = MPI_SUCCESS;
    const int n = file_rw(fh, false, offset, buf, count, datatype, status, &err);
    return err ? err : n;
}
__device__ inline int MPI_File_write_at(MPI_File fh, MPI_Offset offset, const void* buf, int count,
                                        MPI_Datatype datatype, MPI_Status* status) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (fh.amode & MPI_MODE_RDONLY) return MPI_ERR_READ_ONLY;
    if (fh.amode & MPI_MODE_SEQUENTIAL) return MPI_ERR_UNSUPPORTED_OPERATION;
    if (offset < 0) return MPI_ERR_ARG;
    int err = MPI_SUCCESS;
    const int n = file_rw(fh, true, offset, (void*)buf, count, datatype, status, &err);
    return err ? err : n;
}

// Collective forms: every member moves its own part at its own pointer.
__device__ inline int MPI_File_read_all(MPI_File fh, void* buf, int count, MPI_Datatype datatype,
                                        MPI_Status* status) {
    return MPI_File_read(fh, buf, count, datatype, status);
}
__device__ inline int MPI_File_write_all(MPI_File fh, const void* buf, int count, MPI_Datatype datatype,
                                         MPI_Status* status) {
    return MPI_File_write(fh, buf, count, datatype, status);
}

// io.cu:507-559: the shared pointer, advanced atomically (the reference took
// a global mutex around seek + read/write + seek).
static __device__ inline int file_shared_rw(const MPI_File& fh, bool write, void* buf, int count, MPI_Datatype dt,
                                     MPI_Status* status) {
    using namespace gpumpi::rt;
    const int es = type_size(dt);
    if (es <= 0) return MPI_ERR_TYPE;
    if (count < 0) return MPI_ERR_COUNT;
    const int me = comm_rank(comm_get(fh.comm));
    const int ees = type_size(file_view(fh.st)[me].etype);
    const unsigned long long adv = (unsigned long long)count * (unsigned long long)es / (unsigned long long)ees;
    unsigned long long at = 0;
    if (elected()) at = __hip_atomic_fetch_add(&fh.st->shared_pos, adv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    at = uniform_u64(at);
    int err = MPI_SUCCESS;
    const int n = file_rw(fh, write, (MPI_Offset)at, buf, count, dt, status, &err);
    return err ? err : n;
}
__device__ inline int MPI_File_read_shared(MPI_File fh, void* buf, int count, MPI_Datatype datatype,
                                           MPI_Status* status) {
    if (!gpumpi::rt::file_valid(fh)) return MPI_ERR_FILE;
    if (!(fh.amode & MPI_MODE_RDONLY) && !(fh.amode & MPI_MODE_RDWR)) return MPI_ERR_AMODE;
    return file_shared_rw(fh, false, buf, count, datatype, status);
}
__device__ inline int MPI_File_write_shared(MPI_File fh, const void* buf, int count, MPI_Datatype datatype,
                                            MPI_Status* status) {
    if (!gpumpi::rt::file_valid(fh)) return MPI_ERR_FILE;
    if (fh.amode & MPI_MODE_RDONLY) return MPI_ERR_READ_ONLY;
    return file_shared_rw(fh, true, (void*)buf, count, datatype, status);
}
// Collective in MPI: every member passes the same arguments; rank 0 moves it.
__device__ inline int MPI_File_seek_shared(MPI_File fh, MPI_Offset offset, int whence) {
    using namespace gpumpi::rt;
    if (!file_valid(fh)) return MPI_ERR_FILE;
    if (whence != MPI_SEEK_SET && whence != MPI_SEEK_CUR) return MPI_ERR_UNSUPPORTED_OPERATION;
    MPI_Barrier(fh.comm);
    int rank = 0, err = MPI_SUCCESS;
    MPI_Comm_rank(fh.comm, &rank);
    if (rank == 0) {
        const long long np = (whence == MPI_SEEK_SET ? 0 : (long long)fh.st->shared_pos) + offset;
        if (np < 0) err = MPI_ERR_UNSUPPORTED_OPERATION;
        else if (elected()) fh.st->shared_pos = (unsigned long long)np;
        wave_sync();
    }
    MPI_Bcast(&err, 1, MPI_INT, 0, fh.comm);
    return err;
}
__device__ inline int MPI_File_get_position_shared(MPI_File fh, MPI_Offset* offset) {
    if (!gpumpi::rt::file_valid(fh)) return MPI_ERR_FILE;
    *offset = (MPI_Offset)__hip_atomic_load(&fh.st->shared_pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return MPI_SUCCESS;
}
