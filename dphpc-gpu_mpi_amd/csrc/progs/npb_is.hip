// NPB IS (integer bucket sort) as a rank program on the device MPI API.
//
// Follows toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c: key generation
// (randlc / find_my_seed / create_seq, is.c:408-553), rank() (is.c:631-930:
// bucket count -> Allreduce of bucket sizes -> redistribution plan ->
// Alltoall of counts -> Alltoallv of keys -> local counting rank -> partial
// verification against the class's known answers) and full_verify
// (is.c:561-625, with the neighbour Irecv/Send/Wait).  The MPI calls are the
// benchmark's; the rank-local loops run on every wave of the rank through
// gpumpi_parallel (the keys of one rank are counted, scattered and ranked
// in parallel; order inside a bucket is free, the ranks and the sorted
// output are not).
#include "registry.hpp"

namespace {

// randlc, is.c:408-476: x = a * x mod 2^46 in exact double arithmetic.
__device__ inline double randlc(double* x, double a) {
    const double r23 = 1.0 / 8388608.0, r46 = r23 * r23;
    const double t23 = 8388608.0, t46 = t23 * t23;
    double t1 = r23 * a;
    int j = (int)t1;
    const double a1 = j;
    const double a2 = a - t23 * a1;
    t1 = r23 * *x;
    j = (int)t1;
    const double x1 = j;
    const double x2 = *x - t23 * x1;
    t1 = a1 * x2 + a2 * x1;
    j = (int)(r23 * t1);
    const double t2 = j;
    const double z = t1 - t23 * t2;
    const double t3 = t23 * z + a2 * x2;
    j = (int)(r46 * t3);
    const double t4 = j;
    *x = t3 - t46 * t4;
    return r46 * *x;
}

// x * a^m mod 2^46 by binary powering (the jump of find_my_seed, is.c:489-533).
__device__ inline double lcg_jump(double x, double a, uint64_t m) {
    double an = a;
    while (m) {
        if (m & 1) randlc(&x, an);
        randlc(&an, an);
        m >>= 1;
    }
    return x;
}

// Class table (is.c:60-95) and known answers (is.c:254-274).
struct IsClass {
    int total_keys_log2, max_key_log2, num_buckets_log2;
    int test_index[5], test_rank[5];
};
__device__ inline bool is_class(char c, IsClass* k) {
    switch (c) {
    case 'S': *k = {16, 11, 9, {48427, 17148, 23627, 62548, 4431}, {0, 18, 346, 64917, 65463}}; return true;
    case 'W': *k = {20, 16, 10, {357773, 934767, 875723, 898999, 404505}, {1249, 11698, 1039987, 1043896, 1048018}};
        return true;
    case 'A': *k = {23, 19, 10, {2112377, 662041, 5336171, 3642833, 4250760}, {104, 17523, 123928, 8288932, 8388264}};
        return true;
    case 'B': *k = {25, 21, 10, {41869, 812306, 5102857, 18232239, 26860214}, {33422937, 10244, 59149, 33135281, 99}};
        return true;
    default: return false;
    }
}
// is.c:869-913
__device__ inline long expected_rank(char c, int i, int it, long r) {
    switch (c) {
    case 'S': return i <= 2 ? r + it : r - it;
    case 'W': return i < 2 ? r + it - 2 : r - it;
    case 'A': return i <= 2 ? r + it - 1 : r - (it - 1);
    case 'B': return (i == 1 || i == 2 || i == 4) ? r + it : r - it;
    default: return r;
    }
}

// Exclusive prefix sum of n ints in place (global memory) by one wave of a
// gpumpi_parallel body: lane l scans a contiguous chunk, lanes combine.
__device__ inline int wave_exclusive_scan(int* v, int n, int lane) {
    const int chunk = (n + 63) / 64;
    const int lo = lane * chunk, hi = min(n, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; i++) s += v[i];
    int incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    int run = incl - s;
    for (int i = lo; i < hi; i++) {
        const int x = v[i];
        v[i] = run;
        run += x;
    }
    return __shfl(incl, 63);
}

// out[i] = in[0] + ... + in[i-1] for i < n, by the 64 lanes of one wave:
// lane l sums a contiguous chunk, the lanes scan their sums, then each writes
// its chunk.  Reads only `in`, so the lanes exchange nothing through memory.
__device__ inline void wave_exclusive_scan_to(const int* in, int* out, int n, int lane) {
    const int chunk = (n + 63) / 64;
    const int lo = lane * chunk, hi = min(n, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; i++) s += in[i];
    int incl = s;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    int run = incl - s;
    for (int i = lo; i < hi; i++) {
        out[i] = run;
        run += in[i];
    }
}

__device__ inline int wave_incl_scan(int v, int lane) {
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    return v;
}
__device__ inline int wave_sum_int(int v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// Atomic add among the threads of this rank's workgroup only (the ranking
// counts are the rank's own): workgroup scope lets the cache hierarchy
// serve it, where an agent-scope atomic is carried out past the L2.
__device__ __forceinline__ int wg_atomic_add(int* p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// The same on a pointer known to be the rank's LDS (user LDS): ds_add.
typedef GPUMPI_LDS int lds_int;
__device__ __forceinline__ int lds_atomic_add(int* p, int v) {
    return __hip_atomic_fetch_add((lds_int*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// f(key) for every key of keys[0..n): thread t takes the blocks of 16
// consecutive keys b = t, t + nt, ... (a wave reads 4 KiB contiguous per
// block), D blocks per step so that 4*D 16-byte loads per thread are in
// flight together, then the tail one by one.  The same (t, nt) visits the
// same keys in the same order whatever D.  keys is 16-byte aligned.
template <int D = 1, typename F>
__device__ __forceinline__ void for_each_key16(const int* keys, int n, int t, int nt, const F& f) {
    using gpumpi::rt::u32x4;
    const int nblk = n >> 4;
    int b = t;
    for (; b + (D - 1) * nt < nblk; b += D * nt) {
        u32x4 v[D][4];
#pragma unroll
        for (int d = 0; d < D; d++) {
            const u32x4* p = (const u32x4*)(keys + 16 * (b + d * nt));
            v[d][0] = p[0]; v[d][1] = p[1]; v[d][2] = p[2]; v[d][3] = p[3];
        }
#pragma unroll
        for (int d = 0; d < D; d++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                f((int)v[d][q].x); f((int)v[d][q].y); f((int)v[d][q].z); f((int)v[d][q].w);
            }
    }
    for (; b < nblk; b += nt) {
        const u32x4* p = (const u32x4*)(keys + 16 * b);
        const u32x4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
        f((int)v0.x); f((int)v0.y); f((int)v0.z); f((int)v0.w);
        f((int)v1.x); f((int)v1.y); f((int)v1.z); f((int)v1.w);
        f((int)v2.x); f((int)v2.y); f((int)v2.z); f((int)v2.w);
        f((int)v3.x); f((int)v3.y); f((int)v3.z); f((int)v3.w);
    }
    for (int i = (nblk << 4) + t; i < n; i += nt) f(keys[i]);
}

// Inclusive prefix sum dst[i] = src[0] + ... + src[i] for i < n by every
// wave of the rank, in two gpumpi_parallel passes.  Wave w owns the
// contiguous chunk [w*C, (w+1)*C) (C a multiple of 1024); per step a lane
// holds 16 consecutive elements (four 16-byte loads in flight, the wave
// covers 4 KiB).  Pass 1 leaves the chunk sums in wsum[]; pass 2 scans each
// lane's 16 in registers, the lane totals across the wave, and adds the
// chunk's carry (the sum of the lower chunks) and carry0.  src may be LDS
// and dst global memory; src == dst is allowed; both 16-byte aligned.
__device__ inline int scan_chunk(int n, int nw) { return ((n + nw - 1) / nw + 1023) & ~1023; }
__device__ inline void load16(const int* src, int i, int hi, int* v) {
    using gpumpi::rt::u32x4;
    if (i + 16 <= hi) {
        const u32x4* p = (const u32x4*)(src + i);
        const u32x4 a = p[0], b = p[1], c = p[2], d = p[3];
        v[0] = (int)a.x; v[1] = (int)a.y; v[2] = (int)a.z; v[3] = (int)a.w;
        v[4] = (int)b.x; v[5] = (int)b.y; v[6] = (int)b.z; v[7] = (int)b.w;
        v[8] = (int)c.x; v[9] = (int)c.y; v[10] = (int)c.z; v[11] = (int)c.w;
        v[12] = (int)d.x; v[13] = (int)d.y; v[14] = (int)d.z; v[15] = (int)d.w;
    } else {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = i + k < hi ? src[i + k] : 0;
    }
}
__device__ inline void rank_inclusive_scan(const int* src, int* dst, int n, int* wsum, int carry0 = 0) {
    gpumpi_parallel([=] __device__(int t, int nt) {
        const int w = t >> 6, lane = t & 63;
        const int C = scan_chunk(n, nt >> 6);
        const int lo = w * C, hi = min(n, lo + C);
        int s = 0;
        for (int b = lo; b < hi; b += 1024) {
            int v[16];
            load16(src, b + 16 * lane, hi, v);
#pragma unroll
            for (int k = 0; k < 16; k++) s += v[k];
        }
        s = wave_sum_int(s);
        if (lane == 0) wsum[w] = s;
    });
    gpumpi_parallel([=] __device__(int t, int nt) {
        using gpumpi::rt::u32x4;
        const int w = t >> 6, lane = t & 63;
        const int C = scan_chunk(n, nt >> 6);
        const int lo = w * C, hi = min(n, lo + C);
        int carry = carry0 + wave_sum_int(lane < w ? wsum[lane] : 0);
        for (int b = lo; b < hi; b += 1024) {
            const int i0 = b + 16 * lane;
            int v[16];
            load16(src, i0, hi, v);
#pragma unroll
            for (int k = 1; k < 16; k++) v[k] += v[k - 1];
            const int incl = wave_incl_scan(v[15], lane);
            const int add = carry + incl - v[15];
            if (i0 + 16 <= hi) {
                u32x4* q = (u32x4*)(dst + i0);
                q[0] = u32x4{(uint32_t)(v[0] + add), (uint32_t)(v[1] + add), (uint32_t)(v[2] + add), (uint32_t)(v[3] + add)};
                q[1] = u32x4{(uint32_t)(v[4] + add), (uint32_t)(v[5] + add), (uint32_t)(v[6] + add), (uint32_t)(v[7] + add)};
                q[2] = u32x4{(uint32_t)(v[8] + add), (uint32_t)(v[9] + add), (uint32_t)(v[10] + add),
                             (uint32_t)(v[11] + add)};
                q[3] = u32x4{(uint32_t)(v[12] + add), (uint32_t)(v[13] + add), (uint32_t)(v[14] + add),
                             (uint32_t)(v[15] + add)};
            } else {
#pragma unroll
                for (int k = 0; k < 16; k++)
                    if (i0 + k < hi) dst[i0 + k] = v[k] + add;
            }
            carry += __shfl(incl, 63);
        }
    });
}

__device__ inline long wave_sum_long(long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)(uint64_t)v, m);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m);
        v += (long)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

}  // namespace

// Phases of one iteration timed by rank 0 (wall clock, summed over the
// timed iterations) when args[6] is given.
enum IsPhase { PH_HIST, PH_PTRS, PH_SCATTER, PH_ALLREDUCE, PH_PLAN, PH_ALLTOALL, PH_ALLTOALLV, PH_SUMS, PH_COUNT,
               PH_PREFIX, PH_VERIFY, PH_NUM };

// args: [1] class char, [2] out (int64[8]: passed, total keys, errors,
// reserved...), [3] times out (double[4]: total max, Alltoallv max, ...),
// [4] sorted keys out (int32[TOTAL_KEYS]) or 0, [5] iterations (default 10),
// [6] phase times out (double[PH_NUM], rank 0) or 0
GPUMPI_PROGRAM(npb_is) {
    MPI_Init(&argc, &argv);
    int my_rank, np_total;
    MPI_Comm_rank(MPI_COMM_WORLD, &my_rank);
    MPI_Comm_size(MPI_COMM_WORLD, &np_total);
    const char cls = (char)prog_int(1);
    int64_t* out = prog_ptr<int64_t>(2);
    double* tout = prog_ptr<double>(3);
    int32_t* keys_out = prog_ptr<int32_t>(4);
    const int max_iter = prog_int(5) > 0 ? (int)prog_int(5) : 10;
    double* phase_out = prog_ptr<double>(6);
    uint64_t ph_ticks[PH_NUM] = {};
    uint64_t ph_t = 0;
    bool ph_on = false;
    // Closes the running phase (rank 0, timed iterations only).
    auto phase = [&](int k) {
        if (!ph_on) return;
        const uint64_t now = gpumpi_ticks();
        ph_ticks[k] += now - ph_t;
        ph_t = now;
    };
    IsClass K;
    if (!is_class(cls, &K)) return 1;
    MPI_Comm comm_work;
    MPI_Comm_dup(MPI_COMM_WORLD, &comm_work);
    const int comm_size = np_total;  // power of two (checked by the host)
    const int NB = 1 << K.num_buckets_log2;
    const int max_key = 1 << K.max_key_log2;
    const int shift = K.max_key_log2 - K.num_buckets_log2;
    const int num_keys = (int)((1L << K.total_keys_log2) / comm_size);
    const int size_of_buffers = comm_size < 256 ? 3 * num_keys / 2 : comm_size < 512 ? 5 * num_keys / 2
                                                                     : comm_size < 1024 ? 4 * num_keys : 13 * num_keys / 2;
    int* key_array = (int*)gpumpi_malloc(sizeof(int) * (size_t)size_of_buffers);
    int* key_buff1 = (int*)gpumpi_malloc(sizeof(int) * (size_t)size_of_buffers);
    int* key_buff2 = (int*)gpumpi_malloc(sizeof(int) * (size_t)size_of_buffers);
    int* counts = (int*)gpumpi_malloc(sizeof(int) * (size_t)(max_key + 8));  // key_buff_ptr region
    int* bucket_size = (int*)gpumpi_malloc(sizeof(int) * (size_t)(NB + 5));
    int* bucket_totals = (int*)gpumpi_malloc(sizeof(int) * (size_t)(NB + 5));
    int* bucket_ptrs = (int*)gpumpi_malloc(sizeof(int) * (size_t)NB);
    int* ptr1 = (int*)gpumpi_malloc(sizeof(int) * (size_t)(NB + 5));
    int* ptr2 = (int*)gpumpi_malloc(sizeof(int) * (size_t)(NB + 5));
    int* send_count = (int*)gpumpi_malloc(sizeof(int) * 4 * (size_t)comm_size);
    int* recv_count = send_count + comm_size;
    int* send_displ = recv_count + comm_size;
    int* recv_displ = send_displ + comm_size;
    int* scratch = (int*)gpumpi_malloc(sizeof(int) * 2048);  // per-thread partials
    int* hist = gpumpi_user_lds<int>();  // NB bucket counters, then (if room) the ranking counts
    // LDS ints after the bucket counters: the ranking counts of a key range
    // that fits, and, earlier in each iteration, the per-wave sub-histograms.
    const int lcap = (int)((gpumpi::rt::ctx()->lds_bytes - gpumpi::rt::ctx()->off_user) / sizeof(int)) - NB;
    int* sub = hist + NB;
    const bool sub_hist = lcap >= gpumpi::rt::ctx()->waves * NB;

    // create_seq (is.c:535-553) for my slice, in parallel: thread t jumps
    // the generator to its first key.
    {
        const double seed0 = lcg_jump(314159265.00, 1220703125.00,
                                      (uint64_t)my_rank * 4 * (uint64_t)num_keys);
        const int k4 = max_key / 4;
        gpumpi_parallel([=] __device__(int t, int nt) {
            const int chunk = (num_keys + nt - 1) / nt;
            const int lo = t * chunk, hi = min(num_keys, lo + chunk);
            if (lo >= hi) return;
            double s = lcg_jump(seed0, 1220703125.00, 4 * (uint64_t)lo);
            for (int i = lo; i < hi; i++) {
                double x = randlc(&s, 1220703125.00);
                x += randlc(&s, 1220703125.00);
                x += randlc(&s, 1220703125.00);
                x += randlc(&s, 1220703125.00);
                key_array[i] = (int)(k4 * x);
            }
        });
    }

    int passed = 0;
    double t_total0 = 0, t_a2a = 0;
    int total_local_keys = 0, min_key_last = 0;
    for (int it_index = 0; it_index <= max_iter; it_index++) {
        const int iteration = it_index == 0 ? 1 : it_index;
        if (it_index == 1) {
            passed = 0;
            MPI_Barrier(comm_work);
            t_total0 = MPI_Wtime();
            t_a2a = 0;
            ph_on = phase_out != nullptr && my_rank == 0;
        }
        if (ph_on) ph_t = gpumpi_ticks();
        if (my_rank == 0) {  // is.c:648-652
            key_array[iteration] = iteration;
            key_array[iteration + 10] = max_key - iteration;
        }
        // bucket counts (is.c:656-674) in LDS, then bucket pointers and the
        // scatter into buckets (is.c:678-690).  With room in the user LDS
        // every wave counts into its own sub-histogram (atomics contend only
        // within the wave) and scatters from its own offsets: thread t takes
        // the same keys in both passes, so wave w's offsets in bucket b are
        // bucket_ptrs[b] plus the lower waves' counts of b.
        if (sub_hist) {
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < (nt >> 6) * NB; i += nt) sub[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                int* h = sub + (t >> 6) * NB;
                for_each_key16<2>(key_array, num_keys, t, nt, [=](int k) { lds_atomic_add(&h[k >> shift], 1); });
            });
        } else {
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < NB; i += nt) hist[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for_each_key16<2>(key_array, num_keys, t, nt, [=](int k) { lds_atomic_add(&hist[k >> shift], 1); });
            });
        }
        phase(PH_HIST);
        for (int i = 0; i < 5; i++) bucket_size[NB + i] = 0;
        for (int i = 0; i < 5; i++)
            if (K.test_index[i] / num_keys == my_rank) bucket_size[NB + i] = key_array[K.test_index[i] % num_keys];
        gpumpi_parallel([=] __device__(int t, int nt) {
            const int nw = nt >> 6;
            for (int i = t; i < NB; i += nt) {
                int c = 0;
                if (sub_hist)
                    for (int w = 0; w < nw; w++) c += sub[w * NB + i];
                else
                    c = hist[i];
                bucket_size[i] = c;
                bucket_ptrs[i] = c;
            }
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            if (t < 64) wave_exclusive_scan(bucket_ptrs, NB, t);
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            const int nw = nt >> 6;
            for (int i = t; i < NB; i += nt) {
                if (sub_hist) {
                    int run = bucket_ptrs[i];
                    for (int w = 0; w < nw; w++) {
                        const int c = sub[w * NB + i];
                        sub[w * NB + i] = run;
                        run += c;
                    }
                } else {
                    hist[i] = bucket_ptrs[i];
                }
            }
        });
        phase(PH_PTRS);
        gpumpi_parallel([=] __device__(int t, int nt) {
            int* h = sub_hist ? sub + (t >> 6) * NB : hist;
            for_each_key16<2>(key_array, num_keys, t, nt, [=](int k) { key_buff1[lds_atomic_add(&h[k >> shift], 1)] = k; });
        });
        phase(PH_SCATTER);
        // is.c:696-701
        MPI_Allreduce(bucket_size, bucket_totals, NB + 5, MPI_INT, MPI_SUM, comm_work);
        phase(PH_ALLREDUCE);
        // redistribution plan, is.c:706-751 (sequential, as in the benchmark)
        {
            long bucket_sum_accumulator = 0;
            int local_acc = 0, j = 0;
            send_displ[0] = 0;
            ptr1[0] = 0;
            if (NB == 64 * 16) {
                // The same sequential walk over the buckets, with both bucket
                // arrays loaded once into registers (16 per lane) and read
                // back lane by lane (v_readlane), not one global load per
                // bucket; the previous rank's displacement, count and last
                // bucket stay in registers too.
                const int lane = gpumpi::rt::lane_id();
                int bt[16], bs[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    bt[r] = bucket_totals[lane * 16 + r];
                    bs[r] = bucket_size[lane * 16 + r];
                }
                int prev_displ = 0, prev_count = 0, prev_ptr2 = 0;
                for (int l = 0; l < 64; l++) {
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        bucket_sum_accumulator += __builtin_amdgcn_readlane(bt[r], l);
                        local_acc += __builtin_amdgcn_readlane(bs[r], l);
                        if (bucket_sum_accumulator >= (long)(j + 1) * num_keys) {
                            const int i = l * 16 + r;
                            send_count[j] = local_acc;
                            if (j != 0) {
                                prev_displ += prev_count;
                                send_displ[j] = prev_displ;
                                ptr1[j] = prev_ptr2 + 1;
                            }
                            ptr2[j] = i;
                            prev_ptr2 = i;
                            prev_count = local_acc;
                            j++;
                            local_acc = 0;
                        }
                    }
                }
            } else {
                for (int i = 0; i < NB; i++) {
                    bucket_sum_accumulator += bucket_totals[i];
                    local_acc += bucket_size[i];
                    if (bucket_sum_accumulator >= (long)(j + 1) * num_keys) {
                        send_count[j] = local_acc;
                        if (j != 0) {
                            send_displ[j] = send_displ[j - 1] + send_count[j - 1];
                            ptr1[j] = ptr2[j - 1] + 1;
                        }
                        ptr2[j++] = i;
                        local_acc = 0;
                    }
                }
            }
            while (j < comm_size) {
                send_count[j] = 0;
                ptr1[j] = 1;
                j++;
            }
        }
        phase(PH_PLAN);
        // is.c:758-781
        const double a0 = MPI_Wtime();
        MPI_Alltoall(send_count, 1, MPI_INT, recv_count, 1, MPI_INT, comm_work);
        // recv_displ = exclusive prefix of recv_count (is.c:766-768), by the
        // program's wave in one pass instead of a loop of dependent loads.
        wave_exclusive_scan_to(recv_count, recv_displ, comm_size, gpumpi::rt::lane_id());
        gpumpi::rt::wave_sync();  // MPI_Alltoallv's lanes read entries other lanes wrote
        phase(PH_ALLTOALL);
        MPI_Alltoallv(key_buff1, send_count, send_displ, MPI_INT, key_buff2, recv_count, recv_displ, MPI_INT,
                      comm_work);
        t_a2a += MPI_Wtime() - a0;
        phase(PH_ALLTOALLV);
        // local ranking, is.c:787-834
        const int min_key_val = ptr1[my_rank] << shift;
        const int max_key_val = ((ptr2[my_rank] + 1) << shift) - 1;
        const int range = max_key_val - min_key_val + 1;
        // keys of the lower ranks' buckets (m) and of mine (nloc): the same
        // sums, with the lanes taking ranks / buckets in parallel
        long m = 0;
        for (int k = gpumpi::rt::lane_id(); k < my_rank; k += 64)
            for (int i = ptr1[k]; i <= ptr2[k]; i++) m += bucket_totals[i];
        m = wave_sum_long(m);
        long jl = 0;
        for (int i = ptr1[my_rank] + gpumpi::rt::lane_id(); i <= ptr2[my_rank]; i += 64) jl += bucket_totals[i];
        const int nloc = (int)wave_sum_long(jl);
        phase(PH_SUMS);
        // Key counts over my key range: in LDS when the range fits beside the
        // bucket counters (the user LDS the launch reserved), then scanned
        // into `counts`.  A wider range (the edge ranks of the key
        // distribution) is counted in LDS windows of wn keys, one pass over
        // the received keys per window, each window scanned straight into its
        // part of `counts` with the keys below it as carry.  Without room
        // for such windows: atomics on `counts` in global memory.
        int* lcounts = hist + NB;
        const int wn = lcap & ~15;
        const bool counts_in_lds = range <= lcap;
        const bool windowed = !counts_in_lds && wn >= 1024;
        if (counts_in_lds) {
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < range; i += nt) lcounts[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for_each_key16<2>(key_buff2, nloc, t, nt, [=](int k) { lds_atomic_add(&lcounts[k - min_key_val], 1); });
            });
        } else if (windowed) {
            int carry = 0;
            for (int lo = 0; lo < range; lo += wn) {
                const int n = min(wn, range - lo);
                const int base = min_key_val + lo;
                gpumpi_parallel([=] __device__(int t, int nt) {
                    for (int i = t; i < n; i += nt) lcounts[i] = 0;
                });
                gpumpi_parallel([=] __device__(int t, int nt) {
                    for_each_key16<3>(key_buff2, nloc, t, nt, [=](int k) {
                        const unsigned d = (unsigned)(k - base);
                        if (d < (unsigned)n) lds_atomic_add(&lcounts[d], 1);
                    });
                });
                // counts + lo stays 16-byte aligned: wn is a multiple of 16
                rank_inclusive_scan(lcounts, counts + lo, n, scratch, carry);
                carry = counts[lo + n - 1];
            }
        } else {
            // 16-byte zero stores (counts is 256-byte aligned), then the tail
            gpumpi_parallel([=] __device__(int t, int nt) {
                const int n4 = range >> 2;
                const gpumpi::rt::u32x4 z = {0u, 0u, 0u, 0u};
                for (int i = t; i < n4; i += nt) ((gpumpi::rt::u32x4*)counts)[i] = z;
                for (int i = (n4 << 2) + t; i < range; i += nt) counts[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for_each_key16(key_buff2, nloc, t, nt, [=](int k) { wg_atomic_add(&counts[k - min_key_val], 1); });
            });
        }
        phase(PH_COUNT);
        // inclusive prefix over the key range (is.c:829-834), coalesced per
        // wave; from the LDS counts straight into `counts` when they are there
        if (!windowed) rank_inclusive_scan(counts_in_lds ? lcounts : counts, counts, range, scratch);
        phase(PH_PREFIX);
        // partial verification, is.c:840-913
        for (int i = 0; i < 5; i++) {
            const int k = bucket_totals[i + NB];
            if (min_key_val <= k && k <= max_key_val) {
                const long below = k - 1 >= min_key_val ? counts[k - 1 - min_key_val] : 0;
                const long key_rank = below + m;
                if (key_rank == expected_rank(cls, i, iteration, K.test_rank[i])) passed++;
            }
        }
        if (it_index == max_iter) {
            total_local_keys = nloc;
            min_key_last = min_key_val;
        }
        phase(PH_VERIFY);
    }
    if (phase_out && my_rank == 0 && gpumpi_is_leader_lane())
        for (int k = 0; k < PH_NUM; k++) phase_out[k] = (double)ph_ticks[k] * MPI_Wtick();
    const double timecounter = MPI_Wtime() - t_total0;
    double maxtime = 0, maxa2a = 0;
    MPI_Reduce(&timecounter, &maxtime, 1, MPI_DOUBLE, MPI_MAX, 0, comm_work);
    MPI_Reduce(&t_a2a, &maxa2a, 1, MPI_DOUBLE, MPI_MAX, 0, comm_work);

    // full_verify, is.c:561-625
    {
        const int nloc = total_local_keys, mk = min_key_last;
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < nloc; i += nt) {
                const int key = key_buff2[i];
                key_array[wg_atomic_add(&counts[key - mk], -1) - 1] = key;
            }
        });
        int* kp = scratch + 2047;  // the lower neighbour's last key
        *kp = 0;
        MPI_Request request = MPI_REQUEST_NULL;
        const int last_local_key = nloc < 1 ? 0 : nloc - 1;
        if (my_rank > 0) MPI_Irecv(kp, 1, MPI_INT, my_rank - 1, 1000, comm_work, &request);
        if (my_rank < comm_size - 1) MPI_Send(&key_array[last_local_key], 1, MPI_INT, my_rank + 1, 1000, comm_work);
        if (my_rank > 0) MPI_Wait(&request, MPI_STATUS_IGNORE);
        int j = 0;
        if (my_rank > 0 && nloc > 0 && *kp > key_array[0]) j++;
        gpumpi_parallel([=] __device__(int t, int nt) {
            int bad = 0;
            for (int i = 1 + t; i < nloc; i += nt) bad += key_array[i - 1] > key_array[i];
            scratch[t] = bad;
        });
        const int nthreads = gpumpi::rt::ctx()->threads;
        for (int t = 0; t < nthreads; t++) j += scratch[t];
        if (j == 0) passed++;
        int errors = j, errors_all = 0;
        int itemp = passed, passed_all = 0;
        MPI_Reduce(&itemp, &passed_all, 1, MPI_INT, MPI_SUM, 0, comm_work);
        MPI_Reduce(&errors, &errors_all, 1, MPI_INT, MPI_SUM, 0, comm_work);
        // Gather the sorted keys for the host (offset = keys on lower ranks).
        int* all_counts = scratch;
        MPI_Allgather(&nloc, 1, MPI_INT, all_counts, 1, MPI_INT, comm_work);
        long off = 0, total = 0;
        for (int r = 0; r < comm_size; r++) {
            if (r < my_rank) off += all_counts[r];
            total += all_counts[r];
        }
        if (keys_out) {
            int* dst = keys_out + off;
            const int* src = key_array;
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < nloc; i += nt) dst[i] = src[i];
            });
        }
        if (my_rank == 0) {
            out[0] = passed_all;
            out[1] = total;
            out[2] = errors_all;
            tout[0] = maxtime;
            tout[1] = maxa2a;
        }
    }
    MPI_Comm_free(&comm_work);
    MPI_Finalize();
    return 0;
}
