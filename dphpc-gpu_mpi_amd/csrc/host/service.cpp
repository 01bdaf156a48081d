// Host side of the device -> host service channel (gpumpi_rt/world.hpp
// SvcSlot, gpumpi_rt/svc.hpp): the launcher's wait loop calls svc_poll,
// which carries out each rank's posted request -- file I/O for MPI_File_*,
// formatted text for the printf family -- and hands the result back.
//
// Follows gpu_libs/gpu_main/libc_processor.cu:15-84 (the reference's host
// handler of I_FOPEN, I_FREAD, I_FWRITE, I_FSEEK, I_FCLOSE, I_FDELETE) with
// the MPI standard's open semantics instead of its fopen modes: an existing
// file is never truncated, MPI_MODE_EXCL fails only on an existing file.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "gpumpi/host.h"
#include "gpumpi/mpi.h"
#include "host/launcher_internal.hpp"

using namespace gpumpi::rt;

namespace {

// The slots are pinned coherent host memory that running kernels read, so
// they are never freed (see launcher.cpp, host_words): a larger world gets a
// larger array and the old one stays allocated until the process ends.
SvcSlot* g_slots = nullptr;
int g_capacity = 0;
int g_active = 0;  // ranks of the launch being served

std::unordered_map<FILE*, std::string> g_files;  // open host files -> path

FILE* lookup(int64_t handle) {
    FILE* f = (FILE*)(uintptr_t)handle;
    return g_files.count(f) ? f : nullptr;
}

int errno_class(int e) {
    switch (e) {
    case ENOENT: return MPI_ERR_NO_SUCH_FILE;
    case EEXIST: return MPI_ERR_FILE_EXISTS;
    case EACCES:
    case EPERM: return MPI_ERR_ACCESS;
    case EROFS: return MPI_ERR_READ_ONLY;
    case ENOSPC: return MPI_ERR_NO_SPACE;
    case EDQUOT: return MPI_ERR_QUOTA;
    case ENAMETOOLONG: return MPI_ERR_BAD_FILE;
    default: return MPI_ERR_IO;
    }
}

void do_open(SvcSlot& s) {
    const char* path = (const char*)s.data;
    const int amode = s.i[0];
    const bool exists = access(path, F_OK) == 0;
    int err = MPI_SUCCESS;
    if (!exists && ((amode & MPI_MODE_RDONLY) || !(amode & MPI_MODE_CREATE))) err = MPI_ERR_NO_SUCH_FILE;
    else if (exists && (amode & MPI_MODE_EXCL)) err = MPI_ERR_FILE_EXISTS;
    FILE* f = nullptr;
    int64_t size = 0;
    if (err == MPI_SUCCESS) {
        f = std::fopen(path, (amode & MPI_MODE_RDONLY) ? "rb" : exists ? "r+b" : "w+b");
        if (!f) {
            err = errno_class(errno);
        } else {
            g_files[f] = path;
            if (fseeko(f, 0, SEEK_END) == 0) size = (int64_t)ftello(f);
        }
    }
    s.l[0] = (int64_t)(uintptr_t)f;
    s.l[1] = size;
    s.i[1] = err;
}

void do_close(SvcSlot& s) {
    FILE* f = lookup(s.l[0]);
    if (!f) {
        s.i[1] = MPI_ERR_FILE;
        return;
    }
    const std::string path = g_files[f];
    g_files.erase(f);
    int err = std::fclose(f) == 0 ? MPI_SUCCESS : MPI_ERR_IO;
    if (s.i[0] && std::remove(path.c_str()) != 0 && err == MPI_SUCCESS) err = errno_class(errno);
    s.i[1] = err;
}

void do_rw(SvcSlot& s, bool write) {
    FILE* f = lookup(s.l[0]);
    const int64_t n = s.i[0];
    if (!f || n < 0 || n > kSvcDataBytes || s.l[1] < 0) {
        s.i[1] = 0;
        s.i[2] = f ? MPI_ERR_ARG : MPI_ERR_FILE;
        return;
    }
    if (fseeko(f, (off_t)s.l[1], SEEK_SET) != 0) {
        s.i[1] = 0;
        s.i[2] = errno_class(errno);
        return;
    }
    size_t moved = write ? std::fwrite(s.data, 1, (size_t)n, f) : std::fread(s.data, 1, (size_t)n, f);
    int err = MPI_SUCCESS;
    if (std::ferror(f)) {
        err = errno_class(errno);
        std::clearerr(f);
    } else if (write && moved < (size_t)n) {
        err = MPI_ERR_NO_SPACE;
    }
    s.i[1] = (int32_t)moved;
    s.i[2] = err;
}

void do_size(SvcSlot& s) {
    FILE* f = lookup(s.l[0]);
    if (!f) {
        s.i[1] = MPI_ERR_FILE;
        return;
    }
    std::fflush(f);
    s.i[1] = fseeko(f, 0, SEEK_END) == 0 ? MPI_SUCCESS : MPI_ERR_IO;
    s.l[1] = (int64_t)ftello(f);
}

void do_delete(SvcSlot& s) {
    const char* path = (const char*)s.data;
    s.i[1] = std::remove(path) == 0 ? MPI_SUCCESS : errno_class(errno);
}

void do_sync(SvcSlot& s) {
    FILE* f = lookup(s.l[0]);
    s.i[1] = !f ? MPI_ERR_FILE : std::fflush(f) == 0 ? MPI_SUCCESS : MPI_ERR_IO;
}

// printf-family formatting: each conversion of the format is handed to the
// host's snprintf with the argument cast to the type its length modifier
// names, so the text is what the host's printf would print.
struct Arg {
    int kind = 0;
    uint64_t bits = 0;
    const char* str = nullptr;
};

template <typename T>
void append(std::string& out, const std::string& spec, T v) {
    const int n = std::snprintf(nullptr, 0, spec.c_str(), v);
    if (n <= 0) return;
    std::vector<char> buf((size_t)n + 1);
    std::snprintf(buf.data(), buf.size(), spec.c_str(), v);
    out.append(buf.data(), (size_t)n);
}

long long as_signed(const Arg& a) {
    if (a.kind == SVC_ARG_DOUBLE) return (long long)__builtin_bit_cast(double, a.bits);
    return (long long)a.bits;
}
unsigned long long as_unsigned(const Arg& a) {
    if (a.kind == SVC_ARG_DOUBLE) return (unsigned long long)__builtin_bit_cast(double, a.bits);
    return (unsigned long long)a.bits;
}
double as_double(const Arg& a) {
    if (a.kind == SVC_ARG_DOUBLE) return __builtin_bit_cast(double, a.bits);
    if (a.kind == SVC_ARG_INT) return (double)(long long)a.bits;
    return (double)a.bits;
}

std::string format(const SvcSlot& s) {
    std::vector<Arg> args;
    int off = s.i[2];
    for (int k = 0; k < s.i[1]; k++) {
        if (off < 0 || off + (int)sizeof(SvcArg) > kSvcDataBytes) break;
        const SvcArg* a = (const SvcArg*)(s.data + off);
        Arg x;
        x.kind = a->kind;
        x.bits = a->bits;
        if (a->kind == SVC_ARG_STR) {
            if (a->len < 0 || off + (int)sizeof(SvcArg) + a->len >= kSvcDataBytes) break;
            x.str = (const char*)(s.data + off + sizeof(SvcArg));
        }
        args.push_back(x);
        off += (int)sizeof(SvcArg) + ((a->len + 1 + 7) & ~7);
    }
    size_t next = 0;
    auto take = [&]() -> Arg { return next < args.size() ? args[next++] : Arg(); };
    std::string out;
    const char* p = (const char*)s.data;
    while (*p) {
        if (*p != '%') {
            out += *p++;
            continue;
        }
        const char* start = p++;
        if (*p == '%') {
            out += '%';
            p++;
            continue;
        }
        std::string spec = "%";
        while (*p && std::strchr("-+ #0'", *p)) spec += *p++;
        if (*p == '*') {
            spec += std::to_string((int)as_signed(take()));
            p++;
        } else {
            while (std::isdigit((unsigned char)*p)) spec += *p++;
        }
        if (*p == '.') {
            spec += *p++;
            if (*p == '*') {
                spec += std::to_string((int)as_signed(take()));
                p++;
            } else {
                while (std::isdigit((unsigned char)*p)) spec += *p++;
            }
        }
        std::string len;
        while (*p && std::strchr("hlLqjzt", *p)) len += *p++;
        const char conv = *p;
        if (conv) p++;
        switch (conv) {
        case 'd':
        case 'i': {
            const long long v = as_signed(take());
            if (len == "hh") append(out, spec + "d", (int)(signed char)v);
            else if (len == "h") append(out, spec + "d", (int)(short)v);
            else if (len.empty()) append(out, spec + "d", (int)v);
            else append(out, spec + "lld", v);
            break;
        }
        case 'u':
        case 'o':
        case 'x':
        case 'X': {
            const unsigned long long v = as_unsigned(take());
            const std::string c(1, conv);
            if (len == "hh") append(out, spec + c, (unsigned)(unsigned char)v);
            else if (len == "h") append(out, spec + c, (unsigned)(unsigned short)v);
            else if (len.empty()) append(out, spec + c, (unsigned)v);
            else append(out, spec + "ll" + c, v);
            break;
        }
        case 'c': append(out, spec + "c", (int)as_signed(take())); break;
        case 's': {
            const Arg a = take();
            append(out, spec + "s", a.kind == SVC_ARG_STR && a.str ? a.str : "(null)");
            break;
        }
        case 'p': append(out, spec + "p", (void*)(uintptr_t)take().bits); break;
        case 'f':
        case 'F':
        case 'e':
        case 'E':
        case 'g':
        case 'G':
        case 'a':
        case 'A': append(out, spec + std::string(1, conv), as_double(take())); break;
        case 'n': take(); break;  // nothing is written back
        default: out.append(start, (size_t)(p - start)); break;
        }
    }
    return out;
}

void do_format(SvcSlot& s) {
    s.data[kSvcDataBytes - 1] = 0;
    const std::string text = format(s);
    s.i[4] = (int32_t)text.size();
    if (s.i[0] == 1 || s.i[0] == 2) {
        FILE* f = s.i[0] == 2 ? stderr : stdout;
        std::fwrite(text.data(), 1, text.size(), f);
        std::fflush(f);
    } else {
        const int cap = s.i[3] < 0 ? 0 : s.i[3] > kSvcDataBytes ? kSvcDataBytes : s.i[3];
        if (cap > 0) {
            const size_t n = text.size() < (size_t)cap - 1 ? text.size() : (size_t)cap - 1;
            std::memcpy(s.data, text.data(), n);
            s.data[n] = 0;
        }
    }
}

void handle(SvcSlot& s) {
    switch (s.op) {
    case SVC_OPEN: s.data[kSvcDataBytes - 1] = 0; do_open(s); break;
    case SVC_CLOSE: do_close(s); break;
    case SVC_READ: do_rw(s, false); break;
    case SVC_WRITE: do_rw(s, true); break;
    case SVC_SIZE: do_size(s); break;
    case SVC_DELETE: s.data[kSvcDataBytes - 1] = 0; do_delete(s); break;
    case SVC_SYNC: do_sync(s); break;
    case SVC_FORMAT: do_format(s); break;
    default: s.i[1] = MPI_ERR_ARG; s.i[2] = MPI_ERR_ARG; break;
    }
}

}  // namespace

namespace gpumpi {
namespace host {

int svc_prepare(int nranks, SvcSlot** out) {
    if (nranks > g_capacity) {
        SvcSlot* p = nullptr;
        const hipError_t e = hipHostMalloc((void**)&p, (size_t)nranks * sizeof(SvcSlot), hipHostMallocCoherent);
        if (e != hipSuccess) {
            set_err("hipHostMalloc of %d service slots failed: %s", nranks, hipGetErrorString(e));
            return GPUMPI_EHIP;
        }
        g_slots = p;  // the previous array stays allocated (see above)
        g_capacity = nranks;
    }
    for (int r = 0; r < nranks; r++) __atomic_store_n(&g_slots[r].state, (uint32_t)SVC_IDLE, __ATOMIC_RELEASE);
    g_active = nranks;
    *out = g_slots;
    return GPUMPI_OK;
}

int svc_poll() {
    int done = 0;
    for (int r = 0; r < g_active; r++) {
        SvcSlot& s = g_slots[r];
        if (__atomic_load_n(&s.state, __ATOMIC_ACQUIRE) != (uint32_t)SVC_POSTED) continue;
        handle(s);
        __atomic_store_n(&s.state, (uint32_t)SVC_DONE, __ATOMIC_RELEASE);
        done++;
    }
    return done;
}

}  // namespace host
}  // namespace gpumpi

extern "C" int gpumpi_service_poll(void) { return gpumpi::host::svc_poll(); }
