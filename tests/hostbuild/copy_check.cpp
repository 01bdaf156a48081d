// Host build of copy_part (copy.hpp) under AddressSanitizer, run for every
// lane of a 64-lane (or wider) worker set, over random source/destination
// alignments and lengths.  Memory outside the destination range, and
// outside the 16-byte source blocks that overlap the source range, is
// poisoned: the device reads whole aligned 16-byte blocks of the source
// (never beyond a block holding a source byte), and writes exactly the
// destination range.
#include "gpumpi/mpi.h"

#include <sanitizer/asan_interface.h>

#include <random>
#include <vector>

thread_local int gpumpi_host_lane = 0;
thread_local gpumpi_host_idx threadIdx, blockIdx, blockDim, gridDim;
namespace gpumpi {
namespace rt {
thread_local uint8_t (*gpumpi_host_lds)[] = nullptr;
}
}  // namespace gpumpi

int main(int argc, char** argv) {
    const int trials = argc > 1 ? std::atoi(argv[1]) : 2000;
    std::mt19937_64 g(12345);
    long bad = 0;
    for (int tr = 0; tr < trials; tr++) {
        const uint64_t n = (tr % 4 == 0) ? g() % 70000 : g() % 600;
        const int so = (int)(g() % 16), doff = (int)(g() % 16);
        const int nt = (tr % 3 == 0) ? 256 : 64;
        const uint64_t cap = n + 64;
        uint8_t* sbuf = (uint8_t*)aligned_alloc(16, (cap + 15) / 16 * 16);
        uint8_t* dbuf = (uint8_t*)aligned_alloc(16, (cap + 15) / 16 * 16);
        for (uint64_t i = 0; i < cap; i++) {
            sbuf[i] = (uint8_t)g();
            dbuf[i] = 0xA5;
        }
        uint8_t* src = sbuf + 16 + so;
        uint8_t* dst = dbuf + 16 + doff;
        // Poison: source outside the aligned blocks overlapping [src, src+n),
        // destination outside [dst, dst+n).
        const uintptr_t b0 = (uintptr_t)src & ~(uintptr_t)15, b1 = ((uintptr_t)src + n + 15) & ~(uintptr_t)15;
        if (n) {
            __asan_poison_memory_region(sbuf, b0 - (uintptr_t)sbuf);
            __asan_poison_memory_region((void*)b1, (uintptr_t)sbuf + cap - b1);
        }
        __asan_poison_memory_region(dbuf, (uintptr_t)(dst - dbuf));
        __asan_poison_memory_region(dst + n, cap - (uint64_t)(dst + n - dbuf));
        for (int t = 0; t < nt; t++) {
            gpumpi_host_lane = t & 63;
            gpumpi::rt::copy_part(dst, src, n, t, nt);
        }
        __asan_unpoison_memory_region(sbuf, cap);
        __asan_unpoison_memory_region(dbuf, cap);
        if (std::memcmp(dst, src, n) != 0) bad++;
        for (uint64_t i = 0; i < (uint64_t)(dst - dbuf); i++) bad += dbuf[i] != 0xA5;
        for (uint64_t i = (uint64_t)(dst - dbuf) + n; i < cap; i++) bad += dbuf[i] != 0xA5;
        std::free(sbuf);
        std::free(dbuf);
    }
    std::printf("copy trials=%d mismatches=%ld\n", trials, bad);
    return bad ? 1 : 0;
}
