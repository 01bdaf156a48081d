// Resource lifetime inside one launch: the rank heap's free list and
// MPI_Comm_free's return of communicator blocks and context ids.
// Each rank writes ok[world rank] = 1 on success (args[0] = ok).
//   args[1] mode: 0 heap reuse and churn, 1 communicator churn, 2 MPI_Comm_free
//           with requests still pending on the communicator
//   args[2] iterations (modes 0 and 1)
//   args[3] int64[4] out (rank 0): completed iterations, blocks live at the
//           peak, heap bytes at the peak, 0
// The reference's MPI_Comm_free deletes the communicator on its last
// reference (gpu_libs/gpu_mpi/communicator.cu:80-94); its heap is CUDA's
// device malloc/free.
#include "registry.hpp"

namespace {

__device__ inline void set_ok(bool ok) {
    int* okv = prog_ptr<int>(0);
    if (gpumpi_is_leader_lane()) okv[gpumpi_world_rank()] = ok ? 1 : 0;
}

__device__ inline uint32_t lcg(uint32_t& s) {
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

// Fill a block with a pattern of its id, lane-parallel; check it likewise.
__device__ inline void paint(uint8_t* p, uint64_t n, uint32_t id) {
    for (uint64_t i = (uint64_t)gpumpi::rt::lane_id(); i < n; i += 64) p[i] = (uint8_t)(id * 131u + (uint32_t)i);
    gpumpi::rt::wave_sync();
}
__device__ inline bool intact(const uint8_t* p, uint64_t n, uint32_t id) {
    bool bad = false;
    for (uint64_t i = (uint64_t)gpumpi::rt::lane_id(); i < n; i += 64) bad |= p[i] != (uint8_t)(id * 131u + (uint32_t)i);
    return gpumpi::rt::ballot(bad) == 0;
}

__device__ inline bool heap_mode(int iters, int64_t* out) {
    bool ok = true;
    // Reuse, split and merge of free blocks, and the top lowered.
    uint8_t* a = (uint8_t*)gpumpi_malloc(1000);  // 1024
    uint8_t* b = (uint8_t*)gpumpi_malloc(3000);  // 3072
    uint8_t* c = (uint8_t*)gpumpi_malloc(500);   // 512
    ok &= a && b == a + 1024 && c == b + 3072 && ((uintptr_t)a & 255) == 0;
    gpumpi_free(b, 3000);
    uint8_t* d = (uint8_t*)gpumpi_malloc(2000);  // first fit: b's block, split
    uint8_t* e = (uint8_t*)gpumpi_malloc(700);   // the rest of it
    ok &= d == b && e == b + 2048;
    gpumpi_free(a, 1000);
    gpumpi_free(d, 2000);  // merges with a's block below
    uint8_t* f = (uint8_t*)gpumpi_malloc(3072);
    ok &= f == a;
    // c is the top block and merges with the free 256 B left of b's block
    // below it (e took 768 of its 1024), so the top comes down to e's end.
    gpumpi_free(c, 500);
    uint8_t* g = (uint8_t*)gpumpi_malloc(512);
    ok &= g == e + 768;
    gpumpi_free(f, 3072);
    gpumpi_free(g, 512);
    gpumpi_free(e, 700);  // merges both neighbours, ends at the top
    uint8_t* h = (uint8_t*)gpumpi_malloc(1 << 16);
    ok &= h == a;
    gpumpi_free(h, 1 << 16);

    // Churn: up to 16 live blocks of 1 B .. 64 KiB, each painted at
    // allocation and checked before it is freed.
    uint8_t* live[16];
    uint64_t sz[16];
    uint32_t id[16];
    for (int k = 0; k < 16; k++) live[k] = nullptr;
    uint32_t s = 12345u + 7u * (uint32_t)gpumpi_world_rank();
    int peak_n = 0;
    uint64_t bytes = 0, peak_b = 0;
    int it = 0;
    for (; it < iters && ok; it++) {
        const int k = (int)(lcg(s) % 16);
        if (live[k]) {
            ok &= intact(live[k], sz[k], id[k]);
            gpumpi_free(live[k], sz[k]);
            bytes -= sz[k];
            live[k] = nullptr;
        } else {
            const uint32_t r = lcg(s);
            sz[k] = 1 + (r % 4 == 0 ? r % 65536 : r % 2048);
            id[k] = (uint32_t)it;
            live[k] = (uint8_t*)gpumpi_malloc(sz[k]);
            if (!live[k]) {
                ok = false;
                break;
            }
            paint(live[k], sz[k], id[k]);
            bytes += sz[k];
        }
        int n = 0;
        for (int q = 0; q < 16; q++) n += live[q] != nullptr;
        peak_n = n > peak_n ? n : peak_n;
        peak_b = bytes > peak_b ? bytes : peak_b;
    }
    for (int k = 0; k < 16; k++)
        if (live[k]) {
            ok &= intact(live[k], sz[k], id[k]);
            gpumpi_free(live[k], sz[k]);
        }
    // Everything returned: the heap is empty again.
    uint8_t* z = (uint8_t*)gpumpi_malloc(64);
    ok &= z == a;
    gpumpi_free(z, 64);
    if (gpumpi_world_rank() == 0 && gpumpi_is_leader_lane()) {
        out[0] = it;
        out[1] = peak_n;
        out[2] = (int64_t)peak_b;
    }
    return ok;
}

// Communicators created and freed in a loop: MPI_Comm_dup of WORLD every
// iteration, a reversed-key split by parity every 7th, each used by an
// Allreduce before it is freed.  More iterations than there are context
// ids pass only if freed ids come back.
__device__ inline bool comm_mode(int iters, int64_t* out) {
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    int* v = (int*)gpumpi_malloc(2 * sizeof(int));
    bool ok = true;
    int it = 0;
    for (; it < iters && ok; it++) {
        MPI_Comm d = MPI_COMM_NULL;
        ok &= MPI_Comm_dup(MPI_COMM_WORLD, &d) == MPI_SUCCESS && d != MPI_COMM_NULL;
        if (!ok) break;
        if (gpumpi_is_leader_lane()) v[0] = rank + it;
        gpumpi::rt::wave_sync();
        MPI_Allreduce(MPI_IN_PLACE, v, 1, MPI_INT, MPI_SUM, d);
        ok &= v[0] == P * (P - 1) / 2 + P * it;
        MPI_Comm_free(&d);
        ok &= d == MPI_COMM_NULL;
        if (it % 7 == 0) {
            MPI_Comm h;
            MPI_Comm_split(MPI_COMM_WORLD, rank & 1, -rank, &h);
            int hr, hs;
            MPI_Comm_rank(h, &hr);
            MPI_Comm_size(h, &hs);
            if (gpumpi_is_leader_lane()) v[1] = hr;
            gpumpi::rt::wave_sync();
            MPI_Allreduce(MPI_IN_PLACE, v + 1, 1, MPI_INT, MPI_SUM, h);
            ok &= hs == (P - (rank & 1) + 1) / 2 && hr == hs - 1 - rank / 2 && v[1] == hs * (hs - 1) / 2;
            MPI_Comm_free(&h);
        }
    }
    if (rank == 0 && gpumpi_is_leader_lane()) out[0] = it;
    return ok;
}

// Requests still pending when MPI_Comm_free is called complete, and their
// status carries the communicator's rank of the sender (the communicator
// orders WORLD backwards).  Freeing MPI_COMM_WORLD only nulls the handle.
__device__ inline bool pending_mode() {
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    MPI_Comm c;
    MPI_Comm_split(MPI_COMM_WORLD, 0, P - rank, &c);
    int cr;
    MPI_Comm_rank(c, &cr);
    bool ok = cr == P - 1 - rank;
    int* buf = (int*)gpumpi_malloc(64 * sizeof(int));
    for (int i = gpumpi::rt::lane_id(); i < 64; i += 64) buf[i] = -1;
    gpumpi::rt::wave_sync();
    MPI_Request rq = MPI_REQUEST_NULL, sq = MPI_REQUEST_NULL;
    if (cr == 0) {
        MPI_Irecv(buf, 64, MPI_INT, MPI_ANY_SOURCE, 5, c, &rq);
    } else if (cr == 1) {
        for (int i = gpumpi::rt::lane_id(); i < 64; i += 64) buf[i] = 1000 + i;
        gpumpi::rt::wave_sync();
        MPI_Isend(buf, 64, MPI_INT, 0, 5, c, &sq);
        MPI_Request_free(&sq);
    }
    MPI_Comm_free(&c);
    ok &= c == MPI_COMM_NULL;
    if (cr == 0) {
        MPI_Status st;
        MPI_Wait(&rq, &st);
        int n = -1;
        MPI_Get_count(&st, MPI_INT, &n);
        ok &= st.MPI_SOURCE == 1 && st.MPI_TAG == 5 && n == 64;
        bool bad = false;
        for (int i = gpumpi::rt::lane_id(); i < 64; i += 64) bad |= buf[i] != 1000 + i;
        ok &= gpumpi::rt::ballot(bad) == 0;
    }
    MPI_Comm w = MPI_COMM_WORLD;
    ok &= MPI_Comm_free(&w) == MPI_SUCCESS && w == MPI_COMM_NULL;
    // WORLD is still usable
    int one = 1, sum = 0;
    MPI_Allreduce(&one, &sum, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    return ok && sum == P;
}

}  // namespace

GPUMPI_PROGRAM(lifetime) {
    MPI_Init(&argc, &argv);
    const int mode = (int)prog_int(1), iters = (int)prog_int(2);
    int64_t* out = prog_ptr<int64_t>(3);
    bool ok = false;
    if (mode == 0) ok = heap_mode(iters, out);
    else if (mode == 1) ok = comm_mode(iters, out);
    else if (mode == 2) ok = pending_mode();
    MPI_Finalize();
    set_ok(ok);
    return 0;
}
