// Device side of the host service channel: one request at a time per rank
// through the rank's SvcSlot in pinned host memory (world.hpp), carried out
// by the host launcher while the launch runs.  Used by MPI_File_* (io.hpp)
// and by the printf family (include/gpumpi/stdio.h).
//
// Replaces the reference's allocate_host_mem / delegate_to_host / busy wait
// (gpu_libs/gpu_main/device_host_comm.cu:5-15, gpu_libs/gpu_mpi/io.cu:62-148):
// there a managed-memory queue shared by every thread under a lock; here a
// fixed slot per rank, so no lock and no allocation per request.
#pragma once

#include "helpers.hpp"

namespace gpumpi {
namespace rt {

__device__ inline SvcSlot* svc_slot() {
    World* w = world();
    return w->svc ? w->svc + my_rank() : nullptr;
}

// Post the request already written into `s` and wait for the host.  Returns
// false when the launch is aborted first (the slot is then left posted; the
// next launch resets it).
__device__ inline bool svc_call(SvcSlot* s, int op) {
    s->op = op;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (elected()) st_release_system(&s->state, (uint32_t)SVC_POSTED);
    wave_sync();
    Spinner sp;
    while (uniform_u32(ld_acquire_system(&s->state)) != (uint32_t)SVC_DONE) {
        if (!sp.keep_going()) return false;
    }
    return true;
}
// Give the slot back once the results are read.
__device__ inline void svc_release(SvcSlot* s) {
    wave_sync();
    if (elected()) st_release_system(&s->state, (uint32_t)SVC_IDLE);
    wave_sync();
}

// Copy n bytes between device memory and the slot's data, lane-parallel
// (the slot is uncached host memory: 4-byte accesses where both sides are
// aligned, bytes otherwise).  A private (stack) destination is each lane's
// own: there every lane copies all of it.
__device__ inline void svc_copy(void* dst, const void* src, uint64_t n) {
    const int lane = lane_id();
    if (is_private_ptr(dst)) {
        uint8_t* d = (uint8_t*)GPUMPI_LANE_PRIVATE(dst);
        const uint8_t* s = (const uint8_t*)src;
        for (uint64_t i = 0; i < n; i++) d[i] = s[i];
        wave_sync();
        return;
    }
    if ((((uintptr_t)dst | (uintptr_t)src | n) & 3) == 0) {
        uint32_t* d = (uint32_t*)dst;
        const uint32_t* s = (const uint32_t*)src;
        for (uint64_t i = (uint64_t)lane; i < n / 4; i += 64) d[i] = s[i];
    } else {
        uint8_t* d = (uint8_t*)dst;
        const uint8_t* s = (const uint8_t*)src;
        for (uint64_t i = (uint64_t)lane; i < n; i += 64) d[i] = s[i];
    }
    wave_sync();
}

// NUL-terminated string into the slot's data at `off`; returns the bytes
// written including the NUL, or -1 when it does not fit.
__device__ inline int svc_put_string(SvcSlot* s, int off, const char* str) {
    int n = 0;
    if (str)
        while (str[n] != 0) n++;
    if (off + n + 1 > kSvcDataBytes) return -1;
    svc_copy(s->data + off, str ? str : "", (uint64_t)n);
    if (elected()) s->data[off + n] = 0;
    wave_sync();
    return n + 1;
}

// ---------------------------------------------------------------------------
// printf family: the device packs the format and its arguments, the host
// formats them with its own C library (so every conversion behaves exactly
// as the host's printf), and either prints the text or returns it.
// ---------------------------------------------------------------------------
struct FmtArg {
    int32_t kind;
    uint64_t bits;
    const char* str;
};
__device__ inline FmtArg fmt_arg(signed char v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(short v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(int v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(long v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(long long v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(char v) { return {SVC_ARG_INT, (uint64_t)(int64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(bool v) { return {SVC_ARG_INT, (uint64_t)(v ? 1 : 0), nullptr}; }
__device__ inline FmtArg fmt_arg(unsigned char v) { return {SVC_ARG_UINT, (uint64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(unsigned short v) { return {SVC_ARG_UINT, (uint64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(unsigned v) { return {SVC_ARG_UINT, (uint64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(unsigned long v) { return {SVC_ARG_UINT, (uint64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(unsigned long long v) { return {SVC_ARG_UINT, (uint64_t)v, nullptr}; }
__device__ inline FmtArg fmt_arg(float v) {
    const double d = v;
    return {SVC_ARG_DOUBLE, __builtin_bit_cast(uint64_t, d), nullptr};
}
__device__ inline FmtArg fmt_arg(double v) { return {SVC_ARG_DOUBLE, __builtin_bit_cast(uint64_t, v), nullptr}; }
__device__ inline FmtArg fmt_arg(const char* v) { return {SVC_ARG_STR, 0, v}; }
__device__ inline FmtArg fmt_arg(char* v) { return {SVC_ARG_STR, 0, v}; }
template <typename T>
__device__ inline FmtArg fmt_arg(T* v) {
    return {SVC_ARG_PTR, (uint64_t)(uintptr_t)v, nullptr};
}

// Pack and send; stream 0 returns the text into out[0..cap) (NUL-terminated
// when cap > 0).  Returns the formatted length, or -1 without a host or when
// the arguments do not fit one slot.
__device__ inline int svc_format(int stream, char* out, uint64_t cap, const char* fmt, const FmtArg* args,
                                 int nargs) {
    SvcSlot* s = svc_slot();
    if (!s) return -1;
    int off = svc_put_string(s, 0, fmt);
    if (off < 0) return -1;
    off = (off + 7) & ~7;
    const int rec0 = off;
    for (int k = 0; k < nargs; k++) {
        int len = 0;
        if (args[k].kind == SVC_ARG_STR && args[k].str)
            while (args[k].str[len] != 0) len++;
        const int need = (int)sizeof(SvcArg) + ((len + 1 + 7) & ~7);
        if (off + need > kSvcDataBytes) return -1;
        if (elected()) {
            SvcArg* a = (SvcArg*)(s->data + off);
            a->kind = args[k].kind;
            a->len = args[k].kind == SVC_ARG_STR ? len : 0;
            a->bits = args[k].bits;
        }
        if (args[k].kind == SVC_ARG_STR) {
            svc_copy(s->data + off + sizeof(SvcArg), args[k].str ? args[k].str : "", (uint64_t)len);
            if (elected()) s->data[off + (int)sizeof(SvcArg) + len] = 0;
        }
        off += need;
    }
    s->i[0] = stream;
    s->i[1] = nargs;
    s->i[2] = rec0;
    s->i[3] = cap > (uint64_t)kSvcDataBytes ? kSvcDataBytes : (int)cap;
    if (!svc_call(s, SVC_FORMAT)) return -1;
    const int len = s->i[4];
    if (stream == 0 && cap > 0) {
        uint64_t n = cap - 1 < (uint64_t)(kSvcDataBytes - 1) ? cap - 1 : (uint64_t)(kSvcDataBytes - 1);
        if ((uint64_t)len < n) n = (uint64_t)len;
        svc_copy(out, s->data, n);
        if (is_private_ptr(out)) GPUMPI_LANE_PRIVATE(out)[n] = 0;
        else if (elected()) out[n] = 0;
        wave_sync();
    }
    svc_release(s);
    return len;
}

template <typename... A>
__device__ inline int svc_printf(int stream, char* out, uint64_t cap, const char* fmt, A... a) {
    const FmtArg args[sizeof...(A) + 1] = {fmt_arg(a)..., FmtArg{0, 0, nullptr}};
    return svc_format(stream, out, cap, fmt, args, (int)sizeof...(A));
}

}  // namespace rt
}  // namespace gpumpi
