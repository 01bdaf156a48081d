// Point-to-point workloads (ping-pong and ring shift) over the device MPI
// API, following the harness semantics of
// benchmarks/mpi_communication/mpi_communication.c:16-51 (RTT = one
// Send+Recv round trip, max over ranks), plus payload fill/verify kernels.
#include "registry.hpp"

// Payload word i of rank r under seed s: the splitmix64 output function of a
// per-(seed, rank) Weyl sequence (same generator as the CPU oracle).
__device__ __host__ inline uint64_t payload_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __host__ inline uint64_t payload_word(uint64_t seed, uint64_t rank, uint64_t i) {
    return payload_mix64(seed + rank * 0xD1B54A32D192ED03ULL + (i + 1) * 0x9E3779B97F4A7C15ULL);
}

// Buffers of rank r live at base + r * stride.
__global__ void fill_payload_kernel(uint8_t* base, uint64_t stride, uint64_t bytes, uint64_t seed, int first_rank) {
    const int r = first_rank + (int)blockIdx.y;
    uint64_t* p = (uint64_t*)(base + (uint64_t)blockIdx.y * stride);
    const uint64_t nw = bytes / 8;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = payload_word(seed, (uint64_t)r, i);
    if (blockIdx.x == 0 && threadIdx.x < (bytes & 7)) {
        const uint64_t w = payload_word(seed, (uint64_t)r, nw);
        ((uint8_t*)p)[nw * 8 + threadIdx.x] = (uint8_t)(w >> (8 * threadIdx.x));
    }
}

// Counts bytes of buffer j (base + j*stride) that differ from the payload
// of rank src_of(j) = (j + shift) mod nranks.
__global__ void verify_payload_kernel(const uint8_t* base, uint64_t stride, uint64_t bytes, uint64_t seed,
                                      int shift, int nranks, unsigned long long* mismatches) {
    const int j = (int)blockIdx.y;
    const int src = ((j + shift) % nranks + nranks) % nranks;
    const uint64_t* p = (const uint64_t*)(base + (uint64_t)j * stride);
    const uint64_t nw = bytes / 8;
    unsigned long long bad = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw; i += (uint64_t)gridDim.x * blockDim.x)
        bad += p[i] != payload_word(seed, (uint64_t)src, i);
    if (blockIdx.x == 0 && threadIdx.x < (bytes & 7)) {
        const uint64_t w = payload_word(seed, (uint64_t)src, nw);
        bad += ((const uint8_t*)p)[nw * 8 + threadIdx.x] != (uint8_t)(w >> (8 * threadIdx.x));
    }
    if (bad) atomicAdd(mismatches, bad);
}

// Plain device copy (dwordx4 grid-stride), the reference point for HBM
// traffic counters: bytes read and written are exactly n each.
__global__ void copy_calibration_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" {
// Copies `bytes` (a multiple of 16) with copy_calibration_kernel.
int gpumpi_copy_calibration(void* dst, const void* src, uint64_t bytes) {
    if (bytes % 16) return 1;
    hipLaunchKernelGGL(copy_calibration_kernel, dim3(4096), dim3(256), 0, 0, (uint4*)dst, (const uint4*)src,
                       bytes / 16);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
int gpumpi_fill_payload(void* base, uint64_t stride, uint64_t bytes, uint64_t seed, int first_rank, int nbuf) {
    if (nbuf <= 0 || bytes == 0) return 0;
    dim3 grid(256, (unsigned)nbuf);
    hipLaunchKernelGGL(fill_payload_kernel, grid, dim3(256), 0, 0, (uint8_t*)base, stride, bytes, seed, first_rank);
    return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
// Returns mismatching bytes (or -1 on a HIP error).
long long gpumpi_verify_payload(const void* base, uint64_t stride, uint64_t bytes, uint64_t seed, int shift,
                                int nranks, int nbuf) {
    if (nbuf <= 0 || bytes == 0) return 0;
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof *d) != hipSuccess) return -1;
    (void)hipMemset(d, 0, sizeof *d);
    dim3 grid(256, (unsigned)nbuf);
    hipLaunchKernelGGL(verify_payload_kernel, grid, dim3(256), 0, 0, (const uint8_t*)base, stride, bytes, seed, shift,
                       nranks, d);
    unsigned long long h = 0;
    if (hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    (void)hipFree(d);
    return (long long)h;
}
}

// args: [1] sizes ptr (uint64[n]), [2] n sizes, [3] iters ptr (int64[n]),
// [4] warmup ptr (int64[n]), [5] send base, [6] recv base, [7] stride,
// [8] times out (double[n * P]: seconds per rank for the timed iterations)
GPUMPI_PROGRAM(p2p_ring_sweep) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const uint64_t* sizes = prog_ptr<uint64_t>(1);
    const int n = (int)prog_int(2);
    const int64_t* iters = prog_ptr<int64_t>(3);
    const int64_t* warm = prog_ptr<int64_t>(4);
    uint8_t* sbuf = prog_ptr<uint8_t>(5) + (uint64_t)rank * prog_args()[7];
    uint8_t* rbuf = prog_ptr<uint8_t>(6) + (uint64_t)rank * prog_args()[7];
    double* times = prog_ptr<double>(8);
    const int right = (rank + 1) % size, left = (rank + size - 1) % size;
    for (int k = 0; k < n; k++) {
        const int bytes = (int)sizes[k];
        const int64_t total = warm[k] + iters[k];
        double t0 = 0;
        for (int64_t it = 0; it < total; it++) {
            if (it == warm[k]) {
                MPI_Barrier(MPI_COMM_WORLD);
                t0 = MPI_Wtime();
            }
            MPI_Request rq[2];
            MPI_Irecv(rbuf, bytes, MPI_BYTE, left, 7, MPI_COMM_WORLD, &rq[0]);
            MPI_Isend(sbuf, bytes, MPI_BYTE, right, 7, MPI_COMM_WORLD, &rq[1]);
            MPI_Waitall(2, rq, MPI_STATUSES_IGNORE);
        }
        const double t1 = MPI_Wtime();
        if (gpumpi_is_leader_lane()) times[(uint64_t)k * size + rank] = t1 - t0;
        MPI_Barrier(MPI_COMM_WORLD);
    }
    MPI_Finalize();
    return 0;
}

// The same ring under its own kernel name, for untimed warm-up launches (a
// profiler then reports the timed launches alone under p2p_ring_sweep).
GPUMPI_PROGRAM(p2p_ring_warmup) { return prog_p2p_ring_sweep(argc, argv); }

// args as p2p_ring_sweep plus [9] active pairs.  Pair (2p, 2p+1) for
// p < pairs does Send/Recv | Recv/Send (mpi_communication.c:26-33).
GPUMPI_PROGRAM(p2p_pingpong_sweep) {
    MPI_Init(&argc, &argv);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const uint64_t* sizes = prog_ptr<uint64_t>(1);
    const int n = (int)prog_int(2);
    const int64_t* iters = prog_ptr<int64_t>(3);
    const int64_t* warm = prog_ptr<int64_t>(4);
    uint8_t* sbuf = prog_ptr<uint8_t>(5) + (uint64_t)rank * prog_args()[7];
    uint8_t* rbuf = prog_ptr<uint8_t>(6) + (uint64_t)rank * prog_args()[7];
    double* times = prog_ptr<double>(8);
    const int pairs = (int)prog_int(9);
    const bool active = rank / 2 < pairs && (rank ^ 1) < size;
    const int peer = rank ^ 1;
    for (int k = 0; k < n; k++) {
        const int bytes = (int)sizes[k];
        const int64_t total = warm[k] + iters[k];
        double t0 = 0;
        MPI_Barrier(MPI_COMM_WORLD);
        if (active) {
            for (int64_t it = 0; it < total; it++) {
                if (it == warm[k]) t0 = MPI_Wtime();
                if ((rank & 1) == 0) {
                    MPI_Send(sbuf, bytes, MPI_BYTE, peer, 10, MPI_COMM_WORLD);
                    MPI_Recv(rbuf, bytes, MPI_BYTE, peer, 10, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                } else {
                    MPI_Recv(rbuf, bytes, MPI_BYTE, peer, 10, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                    MPI_Send(sbuf, bytes, MPI_BYTE, peer, 10, MPI_COMM_WORLD);
                }
            }
        }
        const double t1 = MPI_Wtime();
        if (gpumpi_is_leader_lane()) times[(uint64_t)k * size + rank] = active ? t1 - t0 : 0.0;
        MPI_Barrier(MPI_COMM_WORLD);
    }
    MPI_Finalize();
    return 0;
}
