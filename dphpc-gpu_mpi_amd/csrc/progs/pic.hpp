// The particle-in-cell cycle shared by the rank programs `pic` (progs/pic.hip,
// with the BF16-wire energy Allreduce) and `pic_exact` (progs/pic_exact.hip,
// FP64 energy only): pic_program<kWire>.  The two live in separate modules
// so that `pic_exact` carries no compressed-wire fold.
#ifndef GPUMPI_PROGS_PIC_HPP
#define GPUMPI_PROGS_PIC_HPP
#include "registry.hpp"

namespace pic_detail {

__device__ inline uint64_t pic_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline float pic_unit(uint64_t w) { return (float)(w >> 40) * (1.0f / 16777216.0f); }
__device__ inline uint64_t pic_draw(uint64_t seed, int rank, int64_t i, int k) {
    return pic_mix64(seed + (uint64_t)rank * 0xD1B54A32D192ED03ULL +
                     (uint64_t)(i * 32 + k + 1) * 0x9E3779B97F4A7C15ULL);
}

__device__ constexpr int kDx[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
__device__ constexpr int kDy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};

struct PicCtx {
    int px, py, nx, ny, rx, ry, rank, nt;
    int64_t cap, n;
    uint64_t seed;
    float vth, drift, qm, Lx, Ly;
    float4* a;        // particles (x, y, u, v)
    float4* b;        // the other particle buffer
    float4* send[9];  // per-direction send buffers (cap each)
    uint8_t* dir;     // direction of each particle
    int64_t* cnt;     // [nt * 9] per-thread direction counts, then offsets
    int64_t* tot;     // [9] per-direction totals; [9..17] receive counts
    int* count;       // ny x nx
    float* rho;       // (ny+4) x (nx+4)
    float* phi;
    float* ex;        // ny x nx
    float* ey;
    float* strip_s;   // halo strips, 2 * max(nx, ny) floats each
    float* strip_r;
    double* eblk;     // per-1024-particle-block energy sums
    double* e;        // [0] this rank's energy, [1] global
    float* e32;       // [0] this rank's energy as FP32, [1] BF16-wire global
};

__device__ inline int chunk_of(int64_t n, int nt) { return (int)((n + nt - 1) / nt); }

// Direction index 1..8 of (dx, dy) in row order, 0 for (0, 0).
__device__ inline int dir_index(int dx, int dy) {
    const int k = (dy + 1) * 3 + (dx + 1);  // 0..8 with (0,0) at 4
    return k < 4 ? k + 1 : k == 4 ? 0 : k;
}

__device__ inline void deposit_and_rho(PicCtx* cp, bool energy) {
    gpumpi_parallel([=] __device__(int t, int nt) {
        PicCtx& c = *cp;
        for (int i = t; i < c.nx * c.ny; i += nt) c.count[i] = 0;
    });
    gpumpi_parallel([=] __device__(int t, int nt) {
        PicCtx& c = *cp;
        const int x0 = c.rx * c.nx, y0 = c.ry * c.ny;
        for (int64_t i = t; i < c.n; i += nt) {
            const float4 q = c.a[i];
            const int lx = (int)q.x - x0, ly = (int)q.y - y0;
            __hip_atomic_fetch_add(&c.count[(size_t)ly * c.nx + lx], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (energy) {
            // 1024-particle blocks, each summed in order (orc_pic_energy)
            const int64_t nb = (c.n + 1023) / 1024;
            for (int64_t b = t; b < nb; b += nt) {
                double s = 0.0;
                const int64_t hi = (b + 1) * 1024 < c.n ? (b + 1) * 1024 : c.n;
                for (int64_t i = b * 1024; i < hi; i++) {
                    const float4 q = c.a[i];
                    const double u = q.z, v = q.w;
                    s = s + 0.5 * (u * u + v * v);
                }
                c.eblk[b] = s;
            }
        }
    });
    gpumpi_parallel([=] __device__(int t, int nt) {
        PicCtx& c = *cp;
        const int W = c.nx + 4;
        for (int i = t; i < c.nx * c.ny; i += nt) {
            const int x = i % c.nx, y = i / c.nx;
            c.rho[(size_t)(y + 2) * W + x + 2] = (float)c.count[i];
        }
    });
}

// Two-cell ring of rho from the 8 neighbours (oracle.c pic_halo): in step d
// this rank sends its strip facing +d to the neighbour in direction d and
// receives the strip of the neighbour opposite d into its ring facing -d.
__device__ inline int halo_exchange(PicCtx* cp, MPI_Comm comm) {
    PicCtx& c = *cp;
    const int w = c.nx, h = c.ny, W = w + 4;
    for (int d = 1; d <= 8; d++) {
        const int dx = kDx[d], dy = kDy[d];
        const int sx0 = dx < 0 ? 2 : dx > 0 ? w : 2, nxs = dx == 0 ? w : 2;
        const int sy0 = dy < 0 ? 2 : dy > 0 ? h : 2, nys = dy == 0 ? h : 2;
        const int tx0 = dx < 0 ? w + 2 : dx > 0 ? 0 : 2;
        const int ty0 = dy < 0 ? h + 2 : dy > 0 ? 0 : 2;
        const int n = nxs * nys;
        for (int k = gpumpi::rt::lane_id(); k < n; k += 64)
            c.strip_s[k] = c.rho[(size_t)(sy0 + k / nxs) * W + sx0 + k % nxs];
        gpumpi::rt::wave_sync();
        const int to = ((c.ry + dy + c.py) % c.py) * c.px + (c.rx + dx + c.px) % c.px;
        const int from = ((c.ry - dy + c.py) % c.py) * c.px + (c.rx - dx + c.px) % c.px;
        const int err = MPI_Sendrecv(c.strip_s, n, MPI_FLOAT, to, 32 + d, c.strip_r, n, MPI_FLOAT, from, 32 + d, comm,
                                     MPI_STATUS_IGNORE);
        if (err) return err;
        for (int k = gpumpi::rt::lane_id(); k < n; k += 64)
            c.rho[(size_t)(ty0 + k / nxs) * W + tx0 + k % nxs] = c.strip_r[k];
        gpumpi::rt::wave_sync();
    }
    return MPI_SUCCESS;
}

// phi on the interior and one ring, then E on the interior (oracle.c pic_field).
__device__ inline void field(PicCtx* cp) {
    gpumpi_parallel([=] __device__(int t, int nt) {
        PicCtx& c = *cp;
        const int W = c.nx + 4, w2 = c.nx + 2, h2 = c.ny + 2;
        for (int k = t; k < w2 * h2; k += nt) {
            const int i = 1 + k % w2, j = 1 + k / w2;
            const float* r = c.rho + (size_t)j * W + i;
            c.phi[(size_t)j * W + i] = ((((4.0f * r[0] + r[-1]) + r[1]) + r[-W]) + r[W]) * 0.125f;
        }
    });
    gpumpi_parallel([=] __device__(int t, int nt) {
        PicCtx& c = *cp;
        const int W = c.nx + 4;
        for (int k = t; k < c.nx * c.ny; k += nt) {
            const int i = k % c.nx, j = k / c.nx;
            const float* f = c.phi + (size_t)(j + 2) * W + (i + 2);
            c.ex[k] = (f[-1] - f[1]) * 0.5f;
            c.ey[k] = (f[-W] - f[W]) * 0.5f;
        }
    });
}

__device__ inline uint64_t fnv1a(const void* p, uint64_t n) {
    const uint8_t* b = (const uint8_t*)p;
    uint64_t h = 1469598103934665603ULL;
    for (uint64_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ULL;
    return h;
}

}  // namespace pic_detail


template <bool kWire>
__device__ int pic_program(int argc, char** argv) {
    using namespace pic_detail;
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int px = (int)prog_int(1), py = (int)prog_int(2), nx = (int)prog_int(3), ny = (int)prog_int(4);
    const int64_t np0 = prog_int(5);
    const int cycles = (int)prog_int(6);
    const int64_t cap = prog_int(17);
    if (px * py != P || nx < 2 || ny < 2 || np0 > cap) {
        MPI_Abort(MPI_COMM_WORLD, 99);
        return 1;
    }
    double* eout = prog_ptr<double>(11);
    float* e16out = prog_ptr<float>(12);
    int64_t* cout = prog_ptr<int64_t>(13);
    uint64_t* phash = prog_ptr<uint64_t>(14);
    uint64_t* rhash = prog_ptr<uint64_t>(15);
    double* tout = prog_ptr<double>(16);
    const int nt = gpumpi::rt::ctx()->threads;

    PicCtx* cp = (PicCtx*)gpumpi_malloc(sizeof(PicCtx));
    PicCtx& c = *cp;
    const size_t grid = (size_t)nx * ny, ring = (size_t)(nx + 4) * (ny + 4);
    const int strip = 2 * (nx > ny ? nx : ny);
    if (gpumpi::rt::elected()) {
        c.px = px; c.py = py; c.nx = nx; c.ny = ny;
        c.rx = rank % px; c.ry = rank / px; c.rank = rank; c.nt = nt;
        c.cap = cap; c.n = np0;
        c.seed = (uint64_t)prog_int(7);
        c.vth = __uint_as_float((unsigned)prog_int(8));
        c.drift = __uint_as_float((unsigned)prog_int(9));
        c.qm = __uint_as_float((unsigned)prog_int(10));
        c.Lx = (float)(px * nx);
        c.Ly = (float)(py * ny);
    }
    gpumpi::rt::wave_sync();
    float4* a = (float4*)gpumpi_malloc(sizeof(float4) * (uint64_t)cap);
    float4* b = (float4*)gpumpi_malloc(sizeof(float4) * (uint64_t)cap);
    float4* sendb = (float4*)gpumpi_malloc(sizeof(float4) * (uint64_t)cap);  // 8 lists share one region
    uint8_t* dir = (uint8_t*)gpumpi_malloc((uint64_t)cap);
    int64_t* cnt = (int64_t*)gpumpi_malloc(sizeof(int64_t) * (uint64_t)nt * 9);
    int64_t* tot = (int64_t*)gpumpi_malloc(sizeof(int64_t) * 32);
    int* count = (int*)gpumpi_malloc(sizeof(int) * grid);
    float* rho = (float*)gpumpi_malloc(sizeof(float) * ring);
    float* phi = (float*)gpumpi_malloc(sizeof(float) * ring);
    float* ex = (float*)gpumpi_malloc(sizeof(float) * grid);
    float* ey = (float*)gpumpi_malloc(sizeof(float) * grid);
    float* ss = (float*)gpumpi_malloc(sizeof(float) * (uint64_t)strip * 2);
    float* sr = (float*)gpumpi_malloc(sizeof(float) * (uint64_t)strip * 2);
    double* eblk = (double*)gpumpi_malloc(sizeof(double) * (uint64_t)((cap + 1023) / 1024 + 1));
    double* e = (double*)gpumpi_malloc(sizeof(double) * 2);
    float* e32 = (float*)gpumpi_malloc(sizeof(float) * 2);
    if (!a || !b || !sendb || !dir || !cnt || !tot || !count || !rho || !phi || !ex || !ey || !ss || !sr || !eblk ||
        !e || !e32) {
        MPI_Abort(MPI_COMM_WORLD, 98);
        return 1;
    }
    if (gpumpi::rt::elected()) {
        c.a = a; c.b = b; c.dir = dir; c.cnt = cnt; c.tot = tot; c.count = count;
        c.rho = rho; c.phi = phi; c.ex = ex; c.ey = ey; c.strip_s = ss; c.strip_r = sr;
        c.eblk = eblk; c.e = e; c.e32 = e32;
        for (int d = 0; d < 9; d++) c.send[d] = sendb;  // set per cycle
    }
    gpumpi::rt::wave_sync();
    // the ring starts at zero, as in the oracle, then every cell of it is received
    gpumpi_parallel([=] __device__(int t, int nt2) {
        for (size_t i = t; i < ring; i += nt2) {
            cp->rho[i] = 0.0f;
            cp->phi[i] = 0.0f;
        }
    });
    // init (orc_pic_init_particle)
    gpumpi_parallel([=] __device__(int t, int nt2) {
        PicCtx& k = *cp;
        const float x0 = (float)(k.rx * k.nx), y0 = (float)(k.ry * k.ny);
        for (int64_t i = t; i < k.n; i += nt2) {
            float x = x0 + (float)k.nx * pic_unit(pic_draw(k.seed, k.rank, i, 0));
            float y = y0 + (float)k.ny * pic_unit(pic_draw(k.seed, k.rank, i, 1));
            if (x >= x0 + (float)k.nx) x = x0;
            if (y >= y0 + (float)k.ny) y = y0;
            float gu = 0.0f, gv = 0.0f;
            for (int j = 0; j < 12; j++) gu = gu + pic_unit(pic_draw(k.seed, k.rank, i, 2 + j));
            for (int j = 0; j < 12; j++) gv = gv + pic_unit(pic_draw(k.seed, k.rank, i, 14 + j));
            k.a[i] = make_float4(x, y, k.drift + k.vth * (gu - 6.0f), k.vth * (gv - 6.0f));
        }
    });
    deposit_and_rho(cp, false);
    if (halo_exchange(cp, MPI_COMM_WORLD)) return 1;
    field(cp);

    double tph[6] = {0, 0, 0, 0, 0, 0};
    MPI_Barrier(MPI_COMM_WORLD);
    const double t_start = MPI_Wtime();
    for (int cyc = 0; cyc < cycles; cyc++) {
        double t0 = MPI_Wtime();
        // push and classify; thread t takes a contiguous chunk so that the
        // partition below keeps particle order (stable)
        gpumpi_parallel([=] __device__(int t, int nt2) {
            PicCtx& k = *cp;
            const int C = chunk_of(k.n, nt2);
            int64_t m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int64_t lo = (int64_t)t * C, hi = lo + C < k.n ? lo + C : k.n;
            for (int64_t i = lo; i < hi; i++) {
                float4 q = k.a[i];
                const int x0 = k.rx * k.nx, y0 = k.ry * k.ny;
                const int lx = (int)q.x - x0, ly = (int)q.y - y0;
                const float e1 = k.ex[(size_t)ly * k.nx + lx], e2 = k.ey[(size_t)ly * k.nx + lx];
                const float u = q.z + k.qm * e1;
                const float v = q.w + k.qm * e2;
                float x = q.x + u, y = q.y + v;
                if (x < 0.0f) x = x + k.Lx;
                if (x >= k.Lx) x = x - k.Lx;
                if (y < 0.0f) y = y + k.Ly;
                if (y >= k.Ly) y = y - k.Ly;
                k.a[i] = make_float4(x, y, u, v);
                const int ox = (int)x / k.nx, oy = (int)y / k.ny;
                const int dx = ox == k.rx ? 0 : ox == (k.rx + 1) % k.px ? 1 : -1;
                const int dy = oy == k.ry ? 0 : oy == (k.ry + 1) % k.py ? 1 : -1;
                const int d = dir_index(dx, dy);
                k.dir[i] = (uint8_t)d;
#pragma unroll
                for (int j = 0; j < 9; j++) m[j] += j == d;  // constant indices: m stays in registers
            }
#pragma unroll
            for (int d = 0; d < 9; d++) k.cnt[(int64_t)t * 9 + d] = m[d];
        });
        // exclusive offsets per direction over threads (lane d), totals
        {
            const int lane = gpumpi::rt::lane_id();
            if (lane < 9) {
                int64_t run = 0;
                for (int t = 0; t < nt; t++) {
                    const int64_t x = cnt[(int64_t)t * 9 + lane];
                    cnt[(int64_t)t * 9 + lane] = run;
                    run += x;
                }
                tot[lane] = run;
            }
            gpumpi::rt::wave_sync();
            // the 8 send lists share one region, back to back in d order
            int64_t base = 0;
            for (int d = 1; d <= 8; d++) {
                if (gpumpi::rt::elected()) c.send[d] = sendb + base;
                base += tot[d];
            }
            gpumpi::rt::wave_sync();
        }
        gpumpi_parallel([=] __device__(int t, int nt2) {
            PicCtx& k = *cp;
            const int C = chunk_of(k.n, nt2);
            int64_t o[9];
#pragma unroll
            for (int d = 0; d < 9; d++) o[d] = k.cnt[(int64_t)t * 9 + d];
            const int64_t lo = (int64_t)t * C, hi = lo + C < k.n ? lo + C : k.n;
            for (int64_t i = lo; i < hi; i++) {
                const int d = k.dir[i];
                const float4 q = k.a[i];
                int64_t pos = 0;
#pragma unroll
                for (int j = 0; j < 9; j++) {  // o[d]++ with constant indices only
                    if (j == d) pos = o[j];
                    o[j] += j == d;
                }
                if (d == 0) k.b[pos] = q;
                else k.send[d][pos] = q;
            }
        });
        double t1 = MPI_Wtime();
        // exchange: counts, then particles, appended after the stays in d order
        int64_t n = tot[0];
        for (int d = 1; d <= 8; d++) {
            const int dx = kDx[d], dy = kDy[d];
            const int to = ((c.ry + dy + py) % py) * px + (c.rx + dx + px) % px;
            const int from = ((c.ry - dy + py) % py) * px + (c.rx - dx + px) % px;
            if (MPI_Sendrecv(&tot[d], 1, MPI_LONG_LONG, to, d, &tot[9 + d], 1, MPI_LONG_LONG, from, d, MPI_COMM_WORLD,
                             MPI_STATUS_IGNORE))
                return 1;
            const int64_t got = tot[9 + d];
            if (n + got > cap) {
                MPI_Abort(MPI_COMM_WORLD, 97);
                return 1;
            }
            if (MPI_Sendrecv(c.send[d], (int)(4 * tot[d]), MPI_FLOAT, to, 16 + d, b + n, (int)(4 * got), MPI_FLOAT,
                             from, 16 + d, MPI_COMM_WORLD, MPI_STATUS_IGNORE))
                return 1;
            n += got;
        }
        if (gpumpi::rt::elected()) {
            c.a = b;
            c.b = a;
            c.n = n;
        }
        gpumpi::rt::wave_sync();
        a = c.a;
        b = c.b;
        double t2 = MPI_Wtime();
        // deposit, energy (exact FP64 and the BF16 wire)
        deposit_and_rho(cp, true);
        if (gpumpi::rt::elected()) {
            const int64_t nb = (c.n + 1023) / 1024;
            double s = 0.0;
            for (int64_t k = 0; k < nb; k++) s = k == 0 ? eblk[k] : s + eblk[k];
            e[0] = s;
            e32[0] = (float)s;
        }
        gpumpi::rt::wave_sync();
        const double t2b = MPI_Wtime();
        MPI_Allreduce(&e[0], &e[1], 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
        if constexpr (kWire) GPUMPI_Allreduce_wire(&e32[0], &e32[1], 1, MPI_FLOAT, MPI_SUM, MPI_COMM_WORLD, 1);
        if (rank == 0 && gpumpi::rt::elected()) {
            if (eout) eout[cyc] = e[1];
            if (kWire && e16out) e16out[cyc] = e32[1];
        }
        double t3 = MPI_Wtime();
        if (halo_exchange(cp, MPI_COMM_WORLD)) return 1;
        const double t3b = MPI_Wtime();
        field(cp);
        double t4 = MPI_Wtime();
        tph[0] += t1 - t0;
        tph[1] += t2 - t1;
        tph[2] += t2b - t2;
        tph[3] += t3 - t2b;
        tph[4] += t3b - t3;
        tph[5] += t4 - t3b;
    }
    const double total = MPI_Wtime() - t_start;
    if (rank == 0 && tout && gpumpi::rt::elected()) {
        for (int k = 0; k < 6; k++) tout[k] = tph[k];
        tout[6] = total;
    }
    if (cout && gpumpi::rt::elected()) cout[rank] = c.n;
    if (phash) {
        const uint64_t h = fnv1a(c.a, sizeof(float4) * (uint64_t)c.n);
        if (gpumpi::rt::elected()) phash[rank] = h;
    }
    if (rhash) {
        const uint64_t h = fnv1a(c.rho, sizeof(float) * ring);
        if (gpumpi::rt::elected()) rhash[rank] = h;
    }
    gpumpi::rt::wave_sync();
    MPI_Finalize();
    return 0;
}

#endif  // GPUMPI_PROGS_PIC_HPP
