// gpumpi_run_noalloc: the launch of gpumpi_run (launcher.cpp) with the whole
// world -- channels, staging, rank heaps, communicator tables, error record
// -- in one file-scope device array, `scratch`, so that the call allocates
// nothing: no hipMalloc, no hipFree, no pinned host memory, no stream or
// event objects.  For callers that cannot allocate (tight serving loops,
// harnesses that forbid it).  Differences from gpumpi_run:
//   * the world must fit in kScratchBytes (the call says how much it needs);
//   * every call starts from fresh channel state (nothing carries over);
//   * the launch runs on the null stream and is waited for with
//     hipStreamSynchronize; there is no watchdog (timeout_s is ignored) and
//     kernel_ms is host wall time around launch and wait;
//   * each GPU holds its own copy of `scratch` until the process ends, and
//     every call on one GPU uses it: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "gpumpi/host.h"
#include "gpumpi/mpi.h"
#include "host/launcher_internal.hpp"

using namespace gpumpi::rt;
using gpumpi::host::set_err;

constexpr uint64_t kScratchBytes = 64ull << 20;

// The static world: one per GPU, held until the process ends.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr uint64_t kArgvBytes = 4096;
constexpr int kMaxDevices = 64;

// Device address of `scratch` per GPU (hipGetSymbolAddress with that GPU
// current); the host-side name `scratch` is not a device address.
uint8_t* g_scratch[kMaxDevices];

int scratch_of_current_device(uint8_t** out) {
    int dev = 0;
    GPUMPI_HIPOK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_err("device %d: at most %d devices", dev, kMaxDevices);
        return GPUMPI_EINVAL;
    }
    if (!g_scratch[dev]) {
        void* p = nullptr;
        GPUMPI_HIPOK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (uint8_t*)p;
    }
    *out = g_scratch[dev];
    return GPUMPI_OK;
}

// Offsets of the world's pieces inside `scratch`.
struct Layout {
    uint64_t world, comm_world, comm_self, world_of, rank_of_world, wslots, sslots, slots, chan_head, done_words,
        staging, arena, arena_top, comm_pool, comm_pool_top, next_ctx, rank_ret, rank_stats, saves, members_hbm,
        dev_abort, abort_word, err, argv, total, slots_bytes, save_bytes;
};

Layout layout_for(const Config& c) {
    const uint64_t P = (uint64_t)c.nranks;
    Layout L{};
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = (o + bytes + 255) & ~(uint64_t)255;
        return at;
    };
    L.world = take(sizeof(World));
    L.comm_world = take(sizeof(CommShared));
    L.comm_self = take(P * sizeof(CommShared));
    L.world_of = take(P * sizeof(int));
    L.rank_of_world = take(P * sizeof(int));
    L.wslots = take(P * sizeof(CollSlot));
    L.sslots = take(P * sizeof(CollSlot));
    L.slots_bytes = P * P * (uint64_t)c.chan_slots * (uint64_t)c.slot_bytes;
    L.slots = take(L.slots_bytes);
    L.chan_head = take(P * P * sizeof(uint32_t));
    L.done_words = take(P * (uint64_t)c.max_requests * sizeof(uint32_t));
    L.staging = take(P * c.staging_bytes);
    L.arena = take(P * c.arena_bytes);
    L.arena_top = take(P * sizeof(uint64_t));
    L.comm_pool = take(c.comm_pool_bytes);
    L.comm_pool_top = take(sizeof(uint64_t));
    L.next_ctx = take(sizeof(int));
    L.rank_ret = take(P * sizeof(int));
    L.rank_stats = take(P * 8 * sizeof(uint64_t));
    L.save_bytes = save_block_bytes(c.nranks);
    L.saves = take(P * L.save_bytes);
    L.members_hbm = take(P * 2 * P * sizeof(uint64_t));
    L.dev_abort = take(256);
    L.abort_word = take(64);
    L.err = take(sizeof(ErrorRecord));
    L.argv = take(kArgvBytes);
    L.total = o;
    return L;
}

// Zero n 16-byte words (a fill kernel, not hipMemsetAsync, so a caller's
// graph capture records no memset node).
__global__ void noalloc_clear(uint4* p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
}

// WORLD and SELF communicators, identity rank maps, first context id, and
// the World itself (all else starts at zero).
__global__ void noalloc_init(World w, World* wdst, int* world_of, int* rank_of_world, CollSlot* wslots,
                             CollSlot* sslots) {
    const int P = w.cfg.nranks;
    for (int i = (int)threadIdx.x; i < P; i += (int)blockDim.x) {
        world_of[i] = i;
        rank_of_world[i] = i;
        CommShared* s = w.comm_self + i;
        s->ctx = 2;
        s->size = 1;
        s->kind = 1;
        s->slots = sslots + i;
    }
    if (threadIdx.x == 0) {
        CommShared* c = w.comm_world;
        c->ctx = 1;
        c->size = P;
        c->kind = 0;
        c->world_of = world_of;
        c->rank_of_world = rank_of_world;
        c->slots = wslots;
        *w.next_ctx = 3;
        *wdst = w;
    }
}

}  // namespace

extern "C" int gpumpi_run_noalloc(const gpumpi_config* cfg_in, const void* kernel, int argc, char** argv,
                                  void* user_args, gpumpi_result* res) {
    gpumpi_result local{};
    gpumpi_result& R = res ? *res : local;
    std::memset(&R, 0, sizeof R);
    Config cfg;
    int lds = 0;
    int rc = gpumpi::host::build_config(cfg_in, &cfg, &lds);
    if (rc) return R.status = rc;
    R.lds_bytes = lds;
    const Layout L = layout_for(cfg);
    if (L.total > kScratchBytes) {
        set_err("a world of %d ranks needs %llu bytes; the static world holds %llu (lower num_procs, "
                "staging_bytes, heap_bytes, comm_pool_bytes or max_requests)",
                cfg.nranks, (unsigned long long)L.total, (unsigned long long)kScratchBytes);
        return R.status = GPUMPI_EINVAL;
    }
    // argv image: pointers, then the strings, copied into the argv region.
    uint8_t img[kArgvBytes];
    const uint64_t ptrbytes = (uint64_t)(argc + 1) * sizeof(char*);
    uint64_t need = ptrbytes;
    for (int i = 0; i < argc; i++) need += std::strlen(argv[i]) + 1;
    if (need > kArgvBytes) {
        set_err("argv needs %llu bytes; the static world keeps %llu", (unsigned long long)need,
                (unsigned long long)kArgvBytes);
        return R.status = GPUMPI_EINVAL;
    }
    uint8_t* base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return R.status = rc;
    {
        char** ptrs = (char**)img;
        uint64_t off = ptrbytes;
        for (int i = 0; i < argc; i++) {
            const uint64_t n = std::strlen(argv[i]) + 1;
            ptrs[i] = (char*)(base + L.argv + off);
            std::memcpy(img + off, argv[i], n);
            off += n;
        }
        ptrs[argc] = nullptr;
    }
    if (lds > 48 * 1024)
        GPUMPI_HIPOK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    const int maxr = gpumpi_max_ranks(cfg_in, kernel);
    if (maxr < cfg.nranks) {
        set_err("%d ranks of %d threads cannot all be resident (max %d): grid barriers would deadlock",
                cfg.nranks, cfg.threads, maxr);
        return R.status = GPUMPI_ETOOBIG;
    }
    const hipError_t pend = hipDeviceSynchronize();
    if (pend != hipSuccess) {
        set_err("GPU error pending before this launch: %s", hipGetErrorString(pend));
        return R.status = GPUMPI_EHIP;
    }

    World w{};
    w.cfg = cfg;
    w.slots = base + L.slots;
    w.chan_head = (uint32_t*)(base + L.chan_head);
    w.done_words = (uint32_t*)(base + L.done_words);
    w.staging = base + L.staging;
    w.arena = base + L.arena;
    w.arena_top = (uint64_t*)(base + L.arena_top);
    w.comm_world = (CommShared*)(base + L.comm_world);
    w.comm_self = (CommShared*)(base + L.comm_self);
    w.comm_pool = base + L.comm_pool;
    w.comm_pool_top = (uint64_t*)(base + L.comm_pool_top);
    w.next_ctx = (int*)(base + L.next_ctx);
    w.rank_ret = (int*)(base + L.rank_ret);
    w.rank_stats = (uint64_t*)(base + L.rank_stats);
    w.abort_word = (int*)(base + L.abort_word);
    w.err = (ErrorRecord*)(base + L.err);
    w.user_args = user_args;
    w.saves = base + L.saves;
    w.save_bytes = L.save_bytes;
    w.members_hbm = (uint64_t*)(base + L.members_hbm);
    w.dev_abort = (uint32_t*)(base + L.dev_abort);

    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(noalloc_clear, dim3(1024), dim3(256), 0, 0, (uint4*)base, L.total / 16);
    GPUMPI_HIPOK(hipGetLastError());
    hipLaunchKernelGGL(noalloc_init, dim3(1), dim3(256), 0, 0, w, (World*)(base + L.world),
                       (int*)(base + L.world_of), (int*)(base + L.rank_of_world), (CollSlot*)(base + L.wslots),
                       (CollSlot*)(base + L.sslots));
    GPUMPI_HIPOK(hipGetLastError());
    GPUMPI_HIPOK(hipMemcpy(base + L.argv, img, need, hipMemcpyHostToDevice));
    World* wptr = (World*)(base + L.world);
    char** dargv = (char**)(base + L.argv);
    void* params[] = {(void*)&wptr, (void*)&argc, (void*)&dargv};
    hipError_t le = hipLaunchCooperativeKernel(kernel, dim3((unsigned)cfg.nranks), dim3((unsigned)cfg.threads), params,
                                               (unsigned)lds, 0);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        set_err("cooperative launch of %d x %d failed: %s", cfg.nranks, cfg.threads, hipGetErrorString(le));
        return R.status = (le == hipErrorCooperativeLaunchTooLarge) ? GPUMPI_ETOOBIG : GPUMPI_EHIP;
    }
    const hipError_t ke = hipStreamSynchronize(0);
    R.kernel_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ke != hipSuccess) {
        set_err("kernel failed: %s", hipGetErrorString(ke));
        return R.status = GPUMPI_EHIP;
    }
    // Return codes in fixed-size chunks (no host allocation either).
    int rets[256];
    for (int r0 = 0; r0 < cfg.nranks; r0 += 256) {
        const int n = cfg.nranks - r0 < 256 ? cfg.nranks - r0 : 256;
        GPUMPI_HIPOK(hipMemcpy(rets, base + L.rank_ret + (uint64_t)r0 * sizeof(int), n * sizeof(int),
                               hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++)
            if (rets[i] != 0) R.any_rank_failed = 1;
    }
    ErrorRecord e;
    GPUMPI_HIPOK(hipMemcpy(&e, base + L.err, sizeof e, hipMemcpyDeviceToHost));
    R.mpi_error = e.code;
    R.mpi_error_rank = e.rank;
    R.mpi_error_line = e.line;
    R.mpi_error_count = e.count;
    std::memcpy(R.mpi_error_msg, e.msg, sizeof R.mpi_error_msg);
    R.mpi_error_msg[sizeof R.mpi_error_msg - 1] = 0;
    if (R.mpi_error) {
        set_err("device MPI error %d on rank %d: %s", e.code, e.rank, e.msg);
        return R.status = GPUMPI_EDEVICE;
    }
    return R.status = GPUMPI_OK;
}
