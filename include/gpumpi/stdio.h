/*
 * gpumpi -- printf family for rank programs, on the host service channel.
 *
 * The device packs the format and its arguments into the rank's service
 * slot; the host launcher formats them with its own C library while the
 * launch runs, so every conversion behaves as the host's printf does.  Text
 * goes to the host's stdout or stderr, or back into a device buffer
 * (sprintf, snprintf).  Like every MPI call, these are called by the rank's
 * program with wave-uniform arguments (not from one lane of a wave).
 *
 * Replaces gpu_libs/gpu_libc/stdio.cuh:11-25 and stdarg.cu:6-31 (a device
 * vsnprintf whose text went to CUDA's printf) and the stdio remaps of
 * stdio.h.cuh:20-27.  A converted program defines GPUMPI_STDIO_NAMES before
 * including this header to have printf, fprintf, sprintf, snprintf, puts,
 * fflush, fopen, fclose, FILE, stdout and stderr mean these.
 *
 * A launch without a serving host (gpumpi_run_noalloc) drops the text and
 * returns -1.
 */
#ifndef GPUMPI_STDIO_H
#define GPUMPI_STDIO_H

#include "gpumpi/mpi.h"

typedef struct gpumpi_FILE_s gpumpi_FILE;
#define gpumpi_stdout ((gpumpi_FILE*)1)
#define gpumpi_stderr ((gpumpi_FILE*)2)

template <typename... A>
__device__ inline int gpumpi_printf(const char* fmt, A... a) {
    return gpumpi::rt::svc_printf(1, nullptr, 0, fmt, a...);
}
template <typename... A>
__device__ inline int gpumpi_fprintf(gpumpi_FILE* stream, const char* fmt, A... a) {
    return gpumpi::rt::svc_printf(stream == gpumpi_stderr ? 2 : 1, nullptr, 0, fmt, a...);
}
/* sprintf: the text is at most one service slot (16 KiB) long. */
template <typename... A>
__device__ inline int gpumpi_sprintf(char* out, const char* fmt, A... a) {
    return gpumpi::rt::svc_printf(0, out, (uint64_t)gpumpi::rt::kSvcDataBytes, fmt, a...);
}
template <typename... A>
__device__ inline int gpumpi_snprintf(char* out, size_t n, const char* fmt, A... a) {
    return gpumpi::rt::svc_printf(0, out, (uint64_t)n, fmt, a...);
}
__device__ inline int gpumpi_puts(const char* s) { return gpumpi::rt::svc_printf(1, nullptr, 0, "%s\n", s); }
__device__ inline int gpumpi_fflush(gpumpi_FILE*) { return 0; }

/* fopen for reading ("r", "rb"): the host opens the file read-only, and
 * fopen returns NULL when it cannot (the file does not exist, say).  Other
 * modes return NULL.  fclose closes it on the host.  Reading through the
 * FILE is not provided; the pair tells a program whether a file exists, as
 * NPB's check_timer_flag (common/c_timers.c) asks. */
struct gpumpi_FILE_s {
    int64_t handle;  // the host's open file
    int64_t size;
};
__device__ inline gpumpi_FILE* gpumpi_fopen(const char* path, const char* mode) {
    using namespace gpumpi::rt;
    if (!path || !mode || mode[0] != 'r' || (mode[1] && mode[1] != 'b')) return nullptr;
    SvcSlot* s = svc_slot();
    if (!s || svc_put_string(s, 0, path) < 0) return nullptr;
    s->i[0] = MPI_MODE_RDONLY;
    if (!svc_call(s, SVC_OPEN)) return nullptr;
    const int err = s->i[1];
    const int64_t h = s->l[0], size = s->l[1];
    svc_release(s);
    if (err != MPI_SUCCESS) return nullptr;
    gpumpi_FILE* f = (gpumpi_FILE*)rank_malloc(sizeof(gpumpi_FILE));
    if (f && elected()) {
        f->handle = h;
        f->size = size;
    }
    wave_sync();
    return f;
}
__device__ inline int gpumpi_fclose(gpumpi_FILE* f) {
    using namespace gpumpi::rt;
    if (!f || f == gpumpi_stdout || f == gpumpi_stderr) return 0;
    SvcSlot* s = svc_slot();
    int err = MPI_ERR_IO;
    if (s) {
        s->l[0] = f->handle;
        s->i[0] = 0;
        if (svc_call(s, SVC_CLOSE)) {
            err = s->i[1];
            svc_release(s);
        }
    }
    rank_free(f, sizeof(gpumpi_FILE));
    return err == MPI_SUCCESS ? 0 : -1;
}

#ifdef GPUMPI_STDIO_NAMES
#undef printf
#undef fprintf
#undef sprintf
#undef snprintf
#undef puts
#undef fflush
#undef fopen
#undef fclose
#undef stdout
#undef stderr
#define printf gpumpi_printf
#define fprintf gpumpi_fprintf
#define sprintf gpumpi_sprintf
#define snprintf gpumpi_snprintf
#define puts gpumpi_puts
#define fflush gpumpi_fflush
#define fopen gpumpi_fopen
#define fclose gpumpi_fclose
#define FILE gpumpi_FILE
#define stdout gpumpi_stdout
#define stderr gpumpi_stderr
#endif

#endif /* GPUMPI_STDIO_H */
