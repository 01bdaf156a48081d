// Host build of the all-to-all / all-gather data phase (coll.hpp:
// seg_info, copy_part) under AddressSanitizer.  The
// collective table is filled the way coll_segments publishes it; then every
// member's segment copy runs for each of the 64 lanes of wave 0, so every
// address the device would touch is checked against exactly sized buffers.
// Usage: segments_check <kind 0..3> <P> <blk>   (kinds as SegKind)
#include "gpumpi/mpi.h"

#include <vector>

thread_local int gpumpi_host_lane = 0;
thread_local gpumpi_host_idx threadIdx, blockIdx, blockDim, gridDim;
namespace gpumpi {
namespace rt {
thread_local uint8_t (*gpumpi_host_lds)[] = nullptr;
}
}  // namespace gpumpi

using namespace gpumpi::rt;

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    const int kind = std::atoi(argv[1]), P = std::atoi(argv[2]), blk = std::atoi(argv[3]);
    const int es = 4;
    // Per-member counts (elements): the block test's shapes.
    //   ALLTOALL / ALLGATHER: blk everywhere; ALLGATHERV: member r sends blk + r;
    //   ALLTOALLV: counts[s][d] = (s + 2 * d) % 5 + blk.
    std::vector<int> cfile;
    if (argc > 4) {  // ALLTOALLV counts[src * P + dst] from a file of P*P int32
        FILE* f = std::fopen(argv[4], "rb");
        if (!f) return 2;
        cfile.resize((size_t)P * P);
        if (std::fread(cfile.data(), 4, cfile.size(), f) != cfile.size()) return 2;
        std::fclose(f);
    }
    auto a2av = [&](int s, int d) { return cfile.empty() ? (s + 2 * d) % 5 + blk : cfile[(size_t)s * P + d]; };
    std::vector<int*> sb(P), rb(P), scnt(P), sdsp(P), rcnt(P), rdsp(P);
    std::vector<uint64_t> sbytes(P), rbytes(P);
    std::vector<CollSlot> tab(P);
    for (int r = 0; r < P; r++) {
        uint64_t sn = 0, rn = 0;
        scnt[r] = new int[P];
        sdsp[r] = new int[P];
        rcnt[r] = new int[P];
        rdsp[r] = new int[P];
        for (int k = 0; k < P; k++) {
            int sc = 0, rc = 0;
            switch (kind) {
            case SEG_ALLTOALLV: sc = a2av(r, k); rc = a2av(k, r); break;
            case SEG_ALLTOALL: sc = blk; rc = blk; break;
            case SEG_ALLGATHER: sc = blk; rc = blk; break;
            default: sc = blk + r; rc = blk + k; break;
            }
            scnt[r][k] = sc;
            rcnt[r][k] = rc;
            sdsp[r][k] = k ? sdsp[r][k - 1] + scnt[r][k - 1] : 0;
            rdsp[r][k] = k ? rdsp[r][k - 1] + rcnt[r][k - 1] : 0;
        }
        for (int k = 0; k < P; k++) rn += (uint64_t)rcnt[r][k];
        if (kind == SEG_ALLTOALLV || kind == SEG_ALLTOALL) {
            for (int k = 0; k < P; k++) sn += (uint64_t)scnt[r][k];
        } else {
            sn = (uint64_t)scnt[r][0];
        }
        sbytes[r] = sn * es;
        rbytes[r] = rn * es;
        // Whole 16-byte blocks: copies read aligned 16-byte source blocks
        // (never past a block that holds a source byte).
        sb[r] = (int*)aligned_alloc(16, (sbytes[r] + 16) / 16 * 16);
        rb[r] = (int*)aligned_alloc(16, (rbytes[r] + 16) / 16 * 16);
        for (uint64_t i = 0; i < sn; i++) sb[r][i] = r * 100000 + (int)i;
        std::memset(rb[r], 0xff, rbytes[r]);
        // What coll_segments publishes for member r.
        CollSlot& t = tab[r];
        std::memset(&t, 0, sizeof t);
        t.w[SW_SBUF] = (uint64_t)(uintptr_t)sb[r];
        t.w[SW_RBUF] = (uint64_t)(uintptr_t)rb[r];
        if (kind == SEG_ALLTOALLV) {
            t.w[SW_SCNT] = (uint64_t)(uintptr_t)scnt[r];
            t.w[SW_SDSP] = (uint64_t)(uintptr_t)sdsp[r];
            t.w[SW_RCNT] = (uint64_t)(uintptr_t)rcnt[r];
            t.w[SW_RDSP] = (uint64_t)(uintptr_t)rdsp[r];
        } else if (kind == SEG_ALLTOALL || kind == SEG_ALLGATHER) {
            t.w[SW_SCNT] = (uint64_t)blk;
            t.w[SW_RCNT] = (uint64_t)blk;
        } else {
            t.w[SW_SCNT] = (uint64_t)(blk + r);
            t.w[SW_RCNT] = (uint64_t)(uintptr_t)rcnt[r];
            t.w[SW_RDSP] = (uint64_t)(uintptr_t)rdsp[r];
        }
    }
    for (int me = 0; me < P; me++) {
        Cmd c{};
        c.op = CMD_ALLTOALLV;
        c.i0 = P;
        c.i1 = me;
        c.i2 = kind;
        c.i3 = 0;
        c.i4 = (rbytes[me] / (uint64_t)P) >= (256u << 10) ? 1 : 0;
        c.a[0] = (uint64_t)(uintptr_t)tab.data();
        c.a[1] = es;
        c.a[2] = es;
        // Every segment's descriptor (seg_info) and copy, lane by lane, as
        // one wave does them (run_segments gathers the descriptors across
        // lanes first; that exchange runs in the whole-kernel simulation).
        for (int k = 0; k < P; k++) {
            uint8_t* d;
            const uint8_t* sp;
            uint64_t n;
            bool trunc = false;
            seg_info(c, k, &d, &sp, &n, &trunc);
            if (trunc) std::printf("unexpected truncation at segment %d of rank %d\n", k, me);
            for (int lane = 0; lane < 64; lane++) {
                gpumpi_host_lane = lane;
                copy_part_lanes(d, sp, n, lane, 64);
            }
        }
    }
    // Expected contents.
    long bad = 0;
    for (int r = 0; r < P; r++)
        for (int k = 0; k < P; k++) {
            const int n = rcnt[r][k];
            for (int i = 0; i < n; i++) {
                int want;
                switch (kind) {
                case SEG_ALLTOALLV:
                case SEG_ALLTOALL: want = k * 100000 + sdsp[k][r] + i; break;
                default: want = k * 100000 + i; break;
                }
                if (rb[r][rdsp[r][k] + i] != want) bad++;
            }
        }
    std::printf("kind=%d P=%d blk=%d mismatches=%ld\n", kind, P, blk, bad);
    return bad ? 1 : 0;
}
