// Batched uneven Allreduce: the whole world split into sub-communicators of
// uneven sizes, each reducing an FP32 vector of its own length, in one
// persistent launch -- exact rank-order path or a compressed wire (BF16,
// FP8 e4m3 with per-256 scales) -- plus an on-device FP64 check of every
// communicator's result.
#include "registry.hpp"

namespace {
__device__ __host__ inline uint64_t nmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
}  // namespace

// N(0,1) floats by Box-Muller over counter-based uniforms: element i of
// stream s = f(seed, s, i).  Test data only.
__global__ void fill_normal_kernel(float* p, uint64_t n, uint64_t seed, uint64_t stream) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = nmix64(seed ^ nmix64(stream * 0x9E3779B97F4A7C15ULL + (i >> 1) + 1));
        const float u1 = ((float)(uint32_t)(h >> 40) + 1.0f) * (1.0f / 16777217.0f);
        const float u2 = (float)(uint32_t)((h >> 8) & 0xffffffu) * (1.0f / 16777216.0f);
        const float r = sqrtf(-2.0f * logf(u1));
        p[i] = (i & 1) ? r * sinf(6.283185307179586f * u2) : r * cosf(6.283185307179586f * u2);
    }
}

extern "C" int gpumpi_fill_normal(void* p, uint64_t n, uint64_t seed, uint64_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(fill_normal_kernel, dim3(1024), dim3(256), 0, 0, (float*)p, n, seed, stream);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

// args: [1] colors (int32[P]), [2] counts per colour (int64[C]), [3] in base
// (float), [4] out base (float), [5] element offsets (uint64[P]), [6] wire
// (0 exact, 1 bf16, 2 fp8), [7] iterations, [8] times (double[P]),
// [9] error sums (double[2*C]: sum (y-ref)^2, sum ref^2), [10] FP64 scratch
// base (double, same offsets) or 0 to skip the check
GPUMPI_PROGRAM(batched_allreduce) {
    MPI_Init(&argc, &argv);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    const int* colors = prog_ptr<int>(1);
    const int64_t* counts = prog_ptr<int64_t>(2);
    const uint64_t off = prog_ptr<uint64_t>(5)[rank];
    float* in = prog_ptr<float>(3) + off;
    float* out = prog_ptr<float>(4) + off;
    const int wire = (int)prog_int(6);
    const int iters = (int)prog_int(7);
    double* times = prog_ptr<double>(8);
    double* errs = prog_ptr<double>(9);
    double* ref64 = prog_ptr<double>(10);
    const int color = colors[rank];
    const int n = (int)counts[color];
    MPI_Comm comm;
    MPI_Comm_split(MPI_COMM_WORLD, color, rank, &comm);
    int crank, csize;
    MPI_Comm_rank(comm, &crank);
    MPI_Comm_size(comm, &csize);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    for (int it = 0; it < iters; it++) GPUMPI_Allreduce_wire(in, out, n, MPI_FLOAT, MPI_SUM, comm, wire);
    MPI_Barrier(MPI_COMM_WORLD);
    const double t1 = MPI_Wtime();
    if (times) times[rank] = t1 - t0;
    if (ref64) {
        // FP64 reference of the same inputs (rank-order FP64 fold), then the
        // normwise error of this communicator's result.
        double* ref = ref64 + off;
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < n; i += nt) ref[i] = (double)in[i];
        });
        MPI_Allreduce(MPI_IN_PLACE, ref, n, MPI_DOUBLE, MPI_SUM, comm);
        const int lo = (int)((int64_t)n * crank / csize), hi = (int)((int64_t)n * (crank + 1) / csize);
        double* part = (double*)gpumpi_malloc(sizeof(double) * 2 * 1024);
        gpumpi_parallel([=] __device__(int t, int nt) {
            double s1 = 0, s2 = 0;
            for (int i = lo + t; i < hi; i += nt) {
                const double d = (double)out[i] - ref[i];
                s1 += d * d;
                s2 += ref[i] * ref[i];
            }
            part[2 * t] = s1;
            part[2 * t + 1] = s2;
        });
        double s[2] = {0, 0}, tot[2] = {0, 0};
        const int nthreads = gpumpi::rt::ctx()->threads;
        for (int t = 0; t < nthreads; t++) {
            s[0] += part[2 * t];
            s[1] += part[2 * t + 1];
        }
        MPI_Reduce(s, tot, 2, MPI_DOUBLE, MPI_SUM, 0, comm);
        if (crank == 0) {
            errs[2 * color] = tot[0];
            errs[2 * color + 1] = tot[1];
        }
    }
    MPI_Comm_free(&comm);
    MPI_Finalize();
    return 0;
}
