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

__device__ inline long wave_sum_long(long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)(uint64_t)v, m);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m);
        v += (long)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

}  // namespace

// args: [1] class char, [2] out (int64[8]: passed, total keys, errors,
// reserved...), [3] times out (double[4]: total max, Alltoallv max, ...),
// [4] sorted keys out (int32[TOTAL_KEYS]) or 0, [5] iterations (default 10)
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
        }
        if (my_rank == 0) {  // is.c:648-652
            key_array[iteration] = iteration;
            key_array[iteration + 10] = max_key - iteration;
        }
        // bucket counts (is.c:656-674) in LDS, then bucket pointers and the
        // scatter into buckets (is.c:678-690)
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < NB; i += nt) hist[i] = 0;
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < num_keys; i += nt) atomicAdd(&hist[key_array[i] >> shift], 1);
        });
        for (int i = 0; i < 5; i++) bucket_size[NB + i] = 0;
        for (int i = 0; i < 5; i++)
            if (K.test_index[i] / num_keys == my_rank) bucket_size[NB + i] = key_array[K.test_index[i] % num_keys];
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < NB; i += nt) {
                bucket_size[i] = hist[i];
                bucket_ptrs[i] = hist[i];
            }
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            if (t < 64) wave_exclusive_scan(bucket_ptrs, NB, t);
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < NB; i += nt) hist[i] = bucket_ptrs[i];
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            for (int i = t; i < num_keys; i += nt) {
                const int key = key_array[i];
                key_buff1[atomicAdd(&hist[key >> shift], 1)] = key;
            }
        });
        // is.c:696-701
        MPI_Allreduce(bucket_size, bucket_totals, NB + 5, MPI_INT, MPI_SUM, comm_work);
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
        // is.c:758-781
        const double a0 = MPI_Wtime();
        MPI_Alltoall(send_count, 1, MPI_INT, recv_count, 1, MPI_INT, comm_work);
        // recv_displ = exclusive prefix of recv_count (is.c:766-768), by the
        // program's wave in one pass instead of a loop of dependent loads.
        wave_exclusive_scan_to(recv_count, recv_displ, comm_size, gpumpi::rt::lane_id());
        gpumpi::rt::wave_sync();  // MPI_Alltoallv's lanes read entries other lanes wrote
        MPI_Alltoallv(key_buff1, send_count, send_displ, MPI_INT, key_buff2, recv_count, recv_displ, MPI_INT,
                      comm_work);
        t_a2a += MPI_Wtime() - a0;
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
        // Key counts over my key range: in LDS when the range fits beside the
        // bucket counters (the user LDS the launch reserved), then copied to
        // `counts`; else straight in global memory.
        int* lcounts = hist + NB;
        const int lcap = (int)((gpumpi::rt::ctx()->lds_bytes - gpumpi::rt::ctx()->off_user) / sizeof(int)) - NB;
        if (range <= lcap) {
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < range; i += nt) lcounts[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < nloc; i += nt) atomicAdd(&lcounts[key_buff2[i] - min_key_val], 1);
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < range; i += nt) counts[i] = lcounts[i];
            });
        } else {
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < range; i += nt) counts[i] = 0;
            });
            gpumpi_parallel([=] __device__(int t, int nt) {
                for (int i = t; i < nloc; i += nt) atomicAdd(&counts[key_buff2[i] - min_key_val], 1);
            });
        }
        // inclusive prefix over the key range (is.c:829-834): per-thread
        // chunk sums, a wave scan of the sums, then the chunk scans.
        gpumpi_parallel([=] __device__(int t, int nt) {
            const int chunk = (range + nt - 1) / nt;
            const int lo = t * chunk, hi = min(range, lo + chunk);
            int s = 0;
            for (int i = lo; i < hi; i++) s += counts[i];
            scratch[t] = s;
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            if (t < 64) wave_exclusive_scan(scratch, nt, t);
        });
        gpumpi_parallel([=] __device__(int t, int nt) {
            const int chunk = (range + nt - 1) / nt;
            const int lo = t * chunk, hi = min(range, lo + chunk);
            int run = scratch[t];
            for (int i = lo; i < hi; i++) {
                run += counts[i];
                counts[i] = run;
            }
        });
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
    }
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
                key_array[atomicSub(&counts[key - mk], 1) - 1] = key;
            }
        });
        int k = 0;
        MPI_Request request = MPI_REQUEST_NULL;
        const int last_local_key = nloc < 1 ? 0 : nloc - 1;
        if (my_rank > 0) MPI_Irecv(&k, 1, MPI_INT, my_rank - 1, 1000, comm_work, &request);
        if (my_rank < comm_size - 1) MPI_Send(&key_array[last_local_key], 1, MPI_INT, my_rank + 1, 1000, comm_work);
        if (my_rank > 0) MPI_Wait(&request, MPI_STATUS_IGNORE);
        int j = 0;
        if (my_rank > 0 && nloc > 0 && k > key_array[0]) j++;
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
