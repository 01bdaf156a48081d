// Particle-decomposition PIC: the communication pattern of the reference's
// sputniPIC (toolchain_tests/sputnipic/src/sputniPIC_CPU.c:157-247 with
// src/mpi_comm.c:79-307) on the 2-D cycle of progs/pic.hpp.  Rank 0 holds
// every particle and the whole grid; each cycle
//   MPI_Bcast of the field (ex, ey) from rank 0,
//   MPI_Scatterv of the particles (uneven counts when P does not divide n),
//   push of the rank's share, NGP counts on the whole grid,
//   MPI_Reduce (FP32 SUM) of the counts to rank 0,
//   MPI_Reduce (FP64 SUM) of the kinetic energy,
//   MPI_Gatherv of the pushed particles back to rank 0,
// then rank 0 forms the density with its periodic ring and the field.  The
// cycle is the one the oracle's orc_pic_global restates on the undecomposed
// grid: the particles come back bit for bit (counts are integers, exact in
// FP32), the energy within FP64 summation order.
//
// args: [1..4] px py nx ny (the domain is px*nx x py*ny cells; its particles
// are those of orc_pic_init_particle for the px x py subdomains), [5]
// particles per subdomain, [6] cycles, [7] seed, [8..10] vth drift qm (FP32
// bits), [11] energy out (double[cycles], rank 0), [12] the particles
// (float4[px*py*np], device memory; rank 0 initialises them and holds the
// final state there), [13] phase seconds of rank 0 summed over the cycles
// (double[5]: field broadcast, scatter, push and counts, reductions and
// field, gather) or 0.
#include "pic.hpp"

GPUMPI_PROGRAM(pic_pdecomp) {
    using namespace pic_detail;
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int px = (int)prog_int(1), py = (int)prog_int(2), nx = (int)prog_int(3), ny = (int)prog_int(4);
    const int64_t npd = prog_int(5);
    const int cycles = (int)prog_int(6);
    const int GW = px * nx, GH = py * ny;
    const int64_t n = (int64_t)px * py * npd;
    float4* G = prog_ptr<float4>(12);
    double* eout = prog_ptr<double>(11);
    double* tout = prog_ptr<double>(13);
    if (GW < 2 || GH < 2 || n < 1 || 4 * n > 0x7fffffff || !G) {
        MPI_Abort(MPI_COMM_WORLD, 99);
        return 1;
    }
    // the share of rank r: n / P particles, one more for the first n % P ranks
    int* counts = (int*)gpumpi_malloc(sizeof(int) * 2 * (uint64_t)P);
    int* displs = counts + P;
    for (int r = gpumpi::rt::lane_id(); r < P; r += 64) {
        const int64_t q = n / P, m = n % P;
        counts[r] = (int)(4 * (q + (r < m ? 1 : 0)));
        displs[r] = (int)(4 * (q * r + (r < m ? r : m)));
    }
    gpumpi::rt::wave_sync();
    const int64_t mine = counts[rank] / 4;
    const size_t grid = (size_t)GW * GH, ring = (size_t)(GW + 4) * (GH + 4);
    PicCtx* cp = (PicCtx*)gpumpi_malloc(sizeof(PicCtx));
    float4* L = (float4*)gpumpi_malloc(sizeof(float4) * (uint64_t)(mine > 0 ? mine : 1));
    float* cnt = (float*)gpumpi_malloc(sizeof(float) * grid);
    float* ex = (float*)gpumpi_malloc(sizeof(float) * grid);
    float* ey = (float*)gpumpi_malloc(sizeof(float) * grid);
    double* e = (double*)gpumpi_malloc(sizeof(double) * 66);
    float *sum = nullptr, *rho = nullptr, *phi = nullptr;
    if (rank == 0) {
        sum = (float*)gpumpi_malloc(sizeof(float) * grid);
        rho = (float*)gpumpi_malloc(sizeof(float) * ring);
        phi = (float*)gpumpi_malloc(sizeof(float) * ring);
    }
    if (!counts || !cp || !L || !cnt || !ex || !ey || !e || (rank == 0 && (!sum || !rho || !phi))) {
        MPI_Abort(MPI_COMM_WORLD, 98);
        return 1;
    }
    if (gpumpi::rt::elected()) {
        PicCtx& c = *cp;
        c.px = 1; c.py = 1; c.nx = GW; c.ny = GH; c.rx = 0; c.ry = 0; c.rank = rank;
        c.nt = gpumpi::rt::ctx()->threads;
        c.cap = mine; c.n = mine;
        c.seed = (uint64_t)prog_int(7);
        c.vth = __uint_as_float((unsigned)prog_int(8));
        c.drift = __uint_as_float((unsigned)prog_int(9));
        c.qm = __uint_as_float((unsigned)prog_int(10));
        c.Lx = (float)GW;
        c.Ly = (float)GH;
        c.a = L; c.b = G; c.count = nullptr;
        c.rho = rho; c.phi = phi; c.ex = ex; c.ey = ey;
        c.eblk = e; c.e = e; c.e32 = nullptr;
    }
    gpumpi::rt::wave_sync();
    if (rank == 0) {
        // the particles of subdomain (vx, vy), as orc_pic_init_particle draws them
        gpumpi_parallel([=] __device__(int t, int nt) {
            PicCtx& k = *cp;
            const int64_t N = (int64_t)px * py * npd;
            for (int64_t g = t; g < N; g += nt) {
                const int vr = (int)(g / npd);
                const int64_t i = g % npd;
                const float x0 = (float)((vr % px) * nx), y0 = (float)((vr / px) * ny);
                float x = x0 + (float)nx * pic_unit(pic_draw(k.seed, vr, i, 0));
                float y = y0 + (float)ny * pic_unit(pic_draw(k.seed, vr, i, 1));
                if (x >= x0 + (float)nx) x = x0;
                if (y >= y0 + (float)ny) y = y0;
                float gu = 0.0f, gv = 0.0f;
                for (int j = 0; j < 12; j++) gu = gu + pic_unit(pic_draw(k.seed, vr, i, 2 + j));
                for (int j = 0; j < 12; j++) gv = gv + pic_unit(pic_draw(k.seed, vr, i, 14 + j));
                k.b[g] = make_float4(x, y, k.drift + k.vth * (gu - 6.0f), k.vth * (gv - 6.0f));
            }
        });
    }
    double tph[5] = {0, 0, 0, 0, 0};
    for (int c = -1; c < cycles; c++) {
        const double t0 = MPI_Wtime();
        if (c >= 0) {
            MPI_Bcast(ex, (int)grid, MPI_FLOAT, 0, MPI_COMM_WORLD);
            MPI_Bcast(ey, (int)grid, MPI_FLOAT, 0, MPI_COMM_WORLD);
        }
        const double t1 = MPI_Wtime();
        MPI_Scatterv(G, counts, displs, MPI_FLOAT, L, counts[rank], MPI_FLOAT, 0, MPI_COMM_WORLD);
        const double t2 = MPI_Wtime();
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (size_t i = t; i < grid; i += nt) cnt[i] = 0.0f;
        });
        const bool push = c >= 0;
        gpumpi_parallel([=] __device__(int t, int nt) {
            PicCtx& k = *cp;
            for (int64_t i = t; i < k.n; i += nt) {
                float4 q = k.a[i];
                if (push) {  // oracle.c pic_push on the whole grid
                    const int lx = (int)q.x, ly = (int)q.y;
                    const float e1 = k.ex[(size_t)ly * k.nx + lx], e2 = k.ey[(size_t)ly * k.nx + lx];
                    const float u = q.z + k.qm * e1;
                    const float v = q.w + k.qm * e2;
                    float x = q.x + u, y = q.y + v;
                    if (x < 0.0f) x = x + k.Lx;
                    if (x >= k.Lx) x = x - k.Lx;
                    if (y < 0.0f) y = y + k.Ly;
                    if (y >= k.Ly) y = y - k.Ly;
                    q = make_float4(x, y, u, v);
                    k.a[i] = q;
                }
                __hip_atomic_fetch_add(&cnt[(size_t)(int)q.y * k.nx + (int)q.x], 1.0f, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        });
        double el = 0.0;
        if (push) {
            for (int64_t i = gpumpi::rt::lane_id(); i < mine; i += 64) {
                const double u = L[i].z, v = L[i].w;
                el = el + 0.5 * (u * u + v * v);
            }
            for (int off = 32; off > 0; off >>= 1) el = el + __shfl_xor(el, off);
        }
        const double t3 = MPI_Wtime();
        MPI_Reduce(cnt, sum, (int)grid, MPI_FLOAT, MPI_SUM, 0, MPI_COMM_WORLD);
        if (push) {
            if (gpumpi::rt::elected()) e[0] = el;
            gpumpi::rt::wave_sync();
            MPI_Reduce(&e[0], &e[1], 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
            if (rank == 0 && eout && gpumpi::rt::elected()) eout[c] = e[1];
            gpumpi::rt::wave_sync();
        }
        if (rank == 0) {
            // the density with its two-cell periodic ring, then the field
            gpumpi_parallel([=] __device__(int t, int nt) {
                PicCtx& k = *cp;
                const int W = k.nx + 4;
                for (size_t idx = t; idx < (size_t)(k.ny + 4) * W; idx += nt) {
                    const int i = (int)(idx % W), j = (int)(idx / W);
                    const int gx = (i - 2 + k.nx) % k.nx, gy = (j - 2 + k.ny) % k.ny;
                    k.rho[idx] = sum[(size_t)gy * k.nx + gx];
                }
            });
            field(cp);
        }
        const double t4 = MPI_Wtime();
        if (push) MPI_Gatherv(L, counts[rank], MPI_FLOAT, G, counts, displs, MPI_FLOAT, 0, MPI_COMM_WORLD);
        const double t5 = MPI_Wtime();
        if (c >= 0) {
            tph[0] += t1 - t0;
            tph[1] += t2 - t1;
            tph[2] += t3 - t2;
            tph[3] += t4 - t3;
            tph[4] += t5 - t4;
        }
    }
    if (rank == 0 && tout && gpumpi::rt::elected())
        for (int k = 0; k < 5; k++) tout[k] = tph[k];
    gpumpi::rt::wave_sync();
    MPI_Finalize();
    return 0;
}
