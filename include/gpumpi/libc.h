/*
 * gpumpi -- the C library subset converted rank programs call, for the
 * device.  Like the MPI calls, these run on a rank's program wave with
 * wave-uniform arguments: the heap functions hand every lane the same block
 * of the rank's heap (gpumpi_malloc), and copies split across the lanes
 * unless the destination is private (a lane's own stack).
 *
 * Replaces the subset of gpu_libs/gpu_libc/ (stdlib.cu, string.cu, and the
 * name remaps of stdlib.h.cuh and string.h.cuh) that converted programs use.
 * A converted program defines GPUMPI_LIBC_NAMES before including this header
 * to have the standard names mean these.
 *
 *   atoi atol atoll atof strtol  strlen strcmp strncmp strcpy strncpy strcat
 *   strchr strrchr strstr strdup memcmp memcpy memmove memset  malloc calloc
 *   realloc free  abs labs
 *   rand srand (a per-rank 48-bit LCG, not the host's sequence)
 *   exit (ends the calling rank's program)  getenv (always NULL)
 */
#ifndef GPUMPI_LIBC_H
#define GPUMPI_LIBC_H

#include "gpumpi/mpi.h"

namespace gpumpi {
namespace libc {

__device__ inline bool is_space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
__device__ inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

__device__ inline long long to_ll(const char* s, char** end, int base) {
    const char* p = s;
    while (is_space(*p)) p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    if ((base == 0 || base == 16) && p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) {
        p += 2;
        base = 16;
    } else if (base == 0) {
        base = p[0] == '0' ? 8 : 10;
    }
    unsigned long long v = 0;
    const char* digits0 = p;
    for (;; p++) {
        int d;
        if (is_digit(*p)) d = *p - '0';
        else if (*p >= 'a' && *p <= 'z') d = *p - 'a' + 10;
        else if (*p >= 'A' && *p <= 'Z') d = *p - 'A' + 10;
        else break;
        if (d >= base) break;
        v = v * (unsigned long long)base + (unsigned long long)d;
    }
    if (end) *end = (char*)(p == digits0 ? s : p);
    return neg ? -(long long)v : (long long)v;
}

// Decimal and exponent notation: the value is formed as digits * 10^exp in
// double arithmetic (within an ulp or two of the correctly rounded value).
__device__ inline double to_d(const char* s, char** end) {
    const char* p = s;
    while (is_space(*p)) p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    double v = 0.0;
    int exp10 = 0;
    bool any = false;
    for (; is_digit(*p); p++, any = true) v = v * 10.0 + (*p - '0');
    if (*p == '.') {
        for (p++; is_digit(*p); p++, any = true) {
            v = v * 10.0 + (*p - '0');
            exp10--;
        }
    }
    if (any && (*p == 'e' || *p == 'E')) {
        char* q;
        const long long e = to_ll(p + 1, &q, 10);
        if (q != p + 1) {
            exp10 += (int)e;
            p = q;
        }
    }
    if (end) *end = (char*)(any ? p : s);
    double scale = 1.0, base = 10.0;
    for (int e = exp10 < 0 ? -exp10 : exp10; e; e >>= 1, base *= base)
        if (e & 1) scale *= base;
    v = exp10 < 0 ? v / scale : v * scale;
    return neg ? -v : v;
}

// Heap blocks carry their size in a 16-byte header, so free() needs none.
__device__ inline void* heap_alloc(size_t n) {
    uint8_t* p = (uint8_t*)gpumpi_malloc(n + 16);
    if (!p) return nullptr;
    if (gpumpi::rt::elected()) *(uint64_t*)p = (uint64_t)n;
    gpumpi::rt::wave_sync();
    return p + 16;
}
__device__ inline void heap_free(void* q) {
    if (!q) return;
    uint8_t* p = (uint8_t*)q - 16;
    gpumpi_free(p, (size_t)*(const uint64_t*)p + 16);
}

__device__ inline void* copy(void* dst, const void* src, size_t n) {
    uint8_t* d = (uint8_t*)dst;
    const uint8_t* s = (const uint8_t*)src;
    if (gpumpi::rt::is_private_ptr(dst)) {
        d = GPUMPI_LANE_PRIVATE(d);
        for (size_t i = 0; i < n; i++) d[i] = s[i];
    } else {
        for (size_t i = (size_t)gpumpi::rt::lane_id(); i < n; i += 64) d[i] = s[i];
        gpumpi::rt::wave_sync();
    }
    return dst;
}
__device__ inline void* fill(void* dst, int c, size_t n) {
    uint8_t* d = (uint8_t*)dst;
    if (gpumpi::rt::is_private_ptr(dst)) {
        d = GPUMPI_LANE_PRIVATE(d);
        for (size_t i = 0; i < n; i++) d[i] = (uint8_t)c;
    } else {
        for (size_t i = (size_t)gpumpi::rt::lane_id(); i < n; i += 64) d[i] = (uint8_t)c;
        gpumpi::rt::wave_sync();
    }
    return dst;
}

}  // namespace libc
}  // namespace gpumpi

__device__ inline int gpumpi_c_atoi(const char* s) { return (int)gpumpi::libc::to_ll(s, nullptr, 10); }
__device__ inline long gpumpi_c_atol(const char* s) { return (long)gpumpi::libc::to_ll(s, nullptr, 10); }
__device__ inline long long gpumpi_c_atoll(const char* s) { return gpumpi::libc::to_ll(s, nullptr, 10); }
__device__ inline long gpumpi_c_strtol(const char* s, char** end, int base) {
    return (long)gpumpi::libc::to_ll(s, end, base);
}
__device__ inline double gpumpi_c_atof(const char* s) { return gpumpi::libc::to_d(s, nullptr); }
__device__ inline double gpumpi_c_strtod(const char* s, char** end) { return gpumpi::libc::to_d(s, end); }
__device__ inline size_t gpumpi_c_strlen(const char* s) {
    size_t n = 0;
    while (s[n]) n++;
    return n;
}
__device__ inline int gpumpi_c_strncmp(const char* a, const char* b, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const unsigned char x = (unsigned char)a[i], y = (unsigned char)b[i];
        if (x != y) return x < y ? -1 : 1;
        if (!x) return 0;
    }
    return 0;
}
__device__ inline int gpumpi_c_strcmp(const char* a, const char* b) { return gpumpi_c_strncmp(a, b, (size_t)-1); }
__device__ inline char* gpumpi_c_strcpy(char* d, const char* s) {
    return (char*)gpumpi::libc::copy(d, s, gpumpi_c_strlen(s) + 1);
}
__device__ inline char* gpumpi_c_strncpy(char* d, const char* s, size_t n) {
    const size_t l = gpumpi_c_strlen(s);
    gpumpi::libc::copy(d, s, l < n ? l : n);
    if (l < n) gpumpi::libc::fill(d + l, 0, n - l);
    return d;
}
__device__ inline char* gpumpi_c_strcat(char* d, const char* s) {
    gpumpi_c_strcpy(d + gpumpi_c_strlen(d), s);
    return d;
}
__device__ inline char* gpumpi_c_strchr(const char* s, int c) {
    for (;; s++) {
        if (*s == (char)c) return (char*)s;
        if (!*s) return nullptr;
    }
}
__device__ inline char* gpumpi_c_strrchr(const char* s, int c) {
    const char* last = nullptr;
    for (;; s++) {
        if (*s == (char)c) last = s;
        if (!*s) return (char*)last;
    }
}
__device__ inline char* gpumpi_c_strstr(const char* h, const char* n) {
    if (!*n) return (char*)h;
    for (; *h; h++) {
        size_t i = 0;
        while (n[i] && h[i] == n[i]) i++;
        if (!n[i]) return (char*)h;
    }
    return nullptr;
}
__device__ inline int gpumpi_c_memcmp(const void* a, const void* b, size_t n) {
    const unsigned char *x = (const unsigned char*)a, *y = (const unsigned char*)b;
    for (size_t i = 0; i < n; i++)
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    return 0;
}
__device__ inline void* gpumpi_c_memcpy(void* d, const void* s, size_t n) { return gpumpi::libc::copy(d, s, n); }
// memmove: through a rank-heap buffer when the ranges overlap.
__device__ inline void* gpumpi_c_memmove(void* d, const void* s, size_t n) {
    const uint8_t *a = (const uint8_t*)s, *b = (const uint8_t*)d;
    if (b + n <= a || a + n <= b || n == 0) return gpumpi::libc::copy(d, s, n);
    void* t = gpumpi::libc::heap_alloc(n);
    if (!t) return nullptr;
    gpumpi::libc::copy(t, s, n);
    gpumpi::libc::copy(d, t, n);
    gpumpi::libc::heap_free(t);
    return d;
}
__device__ inline void* gpumpi_c_memset(void* d, int c, size_t n) { return gpumpi::libc::fill(d, c, n); }
__device__ inline void* gpumpi_c_malloc(size_t n) { return gpumpi::libc::heap_alloc(n); }
__device__ inline void* gpumpi_c_calloc(size_t m, size_t n) {
    void* p = gpumpi::libc::heap_alloc(m * n);
    if (p) gpumpi::libc::fill(p, 0, m * n);
    return p;
}
__device__ inline void gpumpi_c_free(void* p) { gpumpi::libc::heap_free(p); }
__device__ inline void* gpumpi_c_realloc(void* p, size_t n) {
    void* q = gpumpi::libc::heap_alloc(n);
    if (p && q) {
        const size_t old = (size_t)*(const uint64_t*)((const uint8_t*)p - 16);
        gpumpi::libc::copy(q, p, old < n ? old : n);
        gpumpi::libc::heap_free(p);
    }
    return q;
}
__device__ inline char* gpumpi_c_strdup(const char* s) {
    const size_t n = gpumpi_c_strlen(s) + 1;
    char* d = (char*)gpumpi::libc::heap_alloc(n);
    if (d) gpumpi::libc::copy(d, s, n);
    return d;
}
// exit(): the calling rank's program ends here with `code` as its return
// value (gpumpi_rank_exit), as the reference's __gpu_exit ended the rank's
// thread (gpu_libs/gpu_libc/stdlib.cu:50-53).  Other ranks are not told:
// programs call MPI_Abort first where they mean the whole job.
__device__ inline void gpumpi_c_exit(int code) { gpumpi_rank_exit(code); }
// getenv(): the device has no environment.
__device__ inline char* gpumpi_c_getenv(const char*) { return nullptr; }
__device__ inline int gpumpi_c_abs(int v) { return v < 0 ? -v : v; }
__device__ inline long gpumpi_c_labs(long v) { return v < 0 ? -v : v; }

// rand(): x = (0x5DEECE66D x + 11) mod 2^48 (drand48's generator), per rank,
// upper 31 bits; seeded by srand as drand48's srand48 does.
#define GPUMPI_C_RAND_MAX 2147483647
static __device__ unsigned long long gpumpi_c_rand_states[GPU_MPI_MAX_RANKS];
__device__ inline unsigned long long* gpumpi_c_rand_state() {
    return &gpumpi_c_rand_states[gpumpi::rt::my_rank() % GPU_MPI_MAX_RANKS];
}
__device__ inline void gpumpi_c_srand(unsigned seed) {
    if (gpumpi::rt::elected()) *gpumpi_c_rand_state() = ((unsigned long long)seed << 16) | 0x330Eull;
    gpumpi::rt::wave_sync();
}
__device__ inline int gpumpi_c_rand() {
    unsigned long long* s = gpumpi_c_rand_state();
    const unsigned long long x = (*s * 0x5DEECE66Dull + 11ull) & ((1ull << 48) - 1);
    gpumpi::rt::wave_sync();
    if (gpumpi::rt::elected()) *s = x;
    gpumpi::rt::wave_sync();
    return (int)(x >> 17);
}

#ifdef GPUMPI_LIBC_NAMES
#define atoi gpumpi_c_atoi
#define atol gpumpi_c_atol
#define atoll gpumpi_c_atoll
#define atof gpumpi_c_atof
#define strtol gpumpi_c_strtol
#define strtod gpumpi_c_strtod
#define strlen gpumpi_c_strlen
#define strcmp gpumpi_c_strcmp
#define strncmp gpumpi_c_strncmp
#define strcpy gpumpi_c_strcpy
#define strncpy gpumpi_c_strncpy
#define strcat gpumpi_c_strcat
#define strchr gpumpi_c_strchr
#define strrchr gpumpi_c_strrchr
#define strstr gpumpi_c_strstr
#define memmove gpumpi_c_memmove
#define memcmp gpumpi_c_memcmp
#define memcpy gpumpi_c_memcpy
#define memset gpumpi_c_memset
#define malloc gpumpi_c_malloc
#define calloc gpumpi_c_calloc
#define realloc gpumpi_c_realloc
#define strdup gpumpi_c_strdup
#define free gpumpi_c_free
#define exit gpumpi_c_exit
#define getenv gpumpi_c_getenv
#define abs gpumpi_c_abs
#define labs gpumpi_c_labs
#define rand gpumpi_c_rand
#define srand gpumpi_c_srand
#undef RAND_MAX
#define RAND_MAX GPUMPI_C_RAND_MAX
#endif

#endif /* GPUMPI_LIBC_H */
