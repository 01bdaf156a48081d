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
 * fflush, stdout and stderr mean these.
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

#ifdef GPUMPI_STDIO_NAMES
#undef printf
#undef fprintf
#undef sprintf
#undef snprintf
#undef puts
#undef fflush
#undef stdout
#undef stderr
#define printf gpumpi_printf
#define fprintf gpumpi_fprintf
#define sprintf gpumpi_sprintf
#define snprintf gpumpi_snprintf
#define puts gpumpi_puts
#define fflush gpumpi_fflush
#define stdout gpumpi_stdout
#define stderr gpumpi_stderr
#endif

#endif /* GPUMPI_STDIO_H */
