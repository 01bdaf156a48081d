// Host side of the runtime: configuration, world allocation, the persistent
// rank-kernel launch and its watchdog.  Built into libgpumpi.so; the C-ABI is
// declared in include/gpumpi/host.h.
//
// Follows the launch contract of gpu_libs/gpu_main/main.cu:133-321: refuse
// rank counts that cannot all be resident (:195-208), run every rank's
// program in one launch, report 1 if any rank's program returned non-zero
// (:127-129), print the device error record (:304-307).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gpumpi/host.h"
#include "gpumpi/mpi.h"
#include "host/launcher_internal.hpp"

using namespace gpumpi::rt;

namespace gpumpi {
namespace host {

thread_local char g_err[512];

void set_err(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

}  // namespace host
}  // namespace gpumpi

using gpumpi::host::g_err;
using gpumpi::host::set_err;

namespace {

#define HIPOK(call)                                                                            \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GPUMPI_EHIP;                                                                \
        }                                                                                      \
    } while (0)

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Device allocations of one world shape, kept across launches of the same
// shape (re-initialised before every launch).
struct WorldAlloc {
    Config cfg{};
    bool valid = false;
    uint8_t* slots = nullptr;
    size_t slots_bytes = 0;
    uint32_t* chan_head = nullptr;
    uint32_t* done_words = nullptr;
    uint8_t* staging = nullptr;
    uint8_t* arena = nullptr;
    uint64_t* arena_top = nullptr;
    CommShared* comm_world = nullptr;
    CommShared* comm_self = nullptr;
    uint8_t* comm_pool = nullptr;
    uint64_t* comm_pool_top = nullptr;
    int* next_ctx = nullptr;
    int* rank_ret = nullptr;
    uint64_t* rank_stats = nullptr;
    int* abort_word = nullptr;  // pinned host
    ErrorRecord* err = nullptr; // pinned host
    World* dworld = nullptr;
    uint8_t* world_tables = nullptr;  // world_of, rank_of_world, slots of WORLD; self tables
    uint8_t* saves = nullptr;         // per-rank channel cursors between launches
    uint64_t* members_hbm = nullptr;  // [P * 2P] collective member tables (large worlds)
    uint32_t* dev_abort = nullptr;    // device abort flag
    uint64_t save_bytes = 0;
    bool need_reset = true;           // next launch starts the channels from zero
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    void release() {
        auto f = [](void* p) { if (p) (void)hipFree(p); };
        f(slots); f(chan_head); f(done_words); f(staging); f(arena); f(arena_top);
        f(comm_world); f(comm_self); f(comm_pool); f(comm_pool_top); f(next_ctx); f(rank_ret);
        f(rank_stats); f(dworld); f(world_tables); f(saves); f(members_hbm); f(dev_abort);
        // abort_word and err are process-lifetime host words (host_words()):
        // never freed here.
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        *this = WorldAlloc();
    }
};

WorldAlloc g_world;

// The abort word and the error record: pinned, coherent host memory, which
// the GPU reads uncached.  They are allocated once per process and never
// freed, because device memory allocated after a free of such memory that a
// kernel has read can later read wrong values on this platform, and every
// world (re)allocation would otherwise free and re-create them right before
// the launch that reads the new world.
struct HostWords {
    int* abort_word = nullptr;   // [0] host abort request, [1] undelivered-message flag
    ErrorRecord* err = nullptr;
};
HostWords g_host_words;

bool same_shape(const Config& a, const Config& b) {
    return a.nranks == b.nranks && a.chan_slots == b.chan_slots && a.slot_bytes == b.slot_bytes &&
           a.max_requests == b.max_requests && a.staging_bytes == b.staging_bytes &&
           a.arena_bytes == b.arena_bytes && a.comm_pool_bytes == b.comm_pool_bytes;
}

int host_words(int** abort_word, ErrorRecord** err) {
    HostWords& H = g_host_words;
    if (!H.abort_word) HIPOK(hipHostMalloc((void**)&H.abort_word, 64, hipHostMallocCoherent));
    if (!H.err) HIPOK(hipHostMalloc((void**)&H.err, sizeof(ErrorRecord), hipHostMallocCoherent));
    *abort_word = H.abort_word;
    *err = H.err;
    return GPUMPI_OK;
}

int alloc_world(const Config& cfg) {
    if (g_world.valid && same_shape(g_world.cfg, cfg)) {
        g_world.cfg = cfg;
        return GPUMPI_OK;
    }
    if (g_world.valid) g_world.release();
    WorldAlloc& W = g_world;
    const uint64_t P = (uint64_t)cfg.nranks;
    W.slots_bytes = P * P * (uint64_t)cfg.chan_slots * (uint64_t)cfg.slot_bytes;
    HIPOK(hipMalloc(&W.slots, W.slots_bytes));
    HIPOK(hipMalloc(&W.chan_head, P * P * sizeof(uint32_t)));
    HIPOK(hipMalloc(&W.done_words, P * (uint64_t)cfg.max_requests * sizeof(uint32_t)));
    HIPOK(hipMalloc(&W.staging, P * cfg.staging_bytes));
    HIPOK(hipMalloc(&W.arena, P * cfg.arena_bytes));
    HIPOK(hipMalloc(&W.arena_top, P * sizeof(uint64_t)));
    HIPOK(hipMalloc(&W.comm_world, sizeof(CommShared)));
    HIPOK(hipMalloc(&W.comm_self, P * sizeof(CommShared)));
    HIPOK(hipMalloc(&W.comm_pool, cfg.comm_pool_bytes));
    HIPOK(hipMalloc(&W.comm_pool_top, sizeof(uint64_t)));
    HIPOK(hipMalloc(&W.next_ctx, sizeof(int)));
    HIPOK(hipMalloc(&W.rank_ret, P * sizeof(int)));
    HIPOK(hipMalloc(&W.rank_stats, P * 8 * sizeof(uint64_t)));
    HIPOK(hipMalloc(&W.dworld, sizeof(World)));
    // WORLD tables: world_of[P], rank_of_world[P] (identity), slots[P];
    // SELF tables: slots[P] (one per rank).
    const uint64_t tables = 2 * P * sizeof(int) + 256 + 2 * P * sizeof(CollSlot);
    HIPOK(hipMalloc(&W.world_tables, tables));
    W.save_bytes = save_block_bytes(cfg.nranks);
    HIPOK(hipMalloc(&W.saves, P * W.save_bytes));
    HIPOK(hipMalloc(&W.members_hbm, P * 2 * P * sizeof(uint64_t)));
    HIPOK(hipMalloc(&W.dev_abort, 256));
    W.need_reset = true;
    {
        const int rc = host_words(&W.abort_word, &W.err);
        if (rc) return rc;
    }
    HIPOK(hipStreamCreateWithFlags(&W.stream, hipStreamNonBlocking));
    HIPOK(hipEventCreate(&W.ev0));
    HIPOK(hipEventCreate(&W.ev1));
    W.cfg = cfg;
    W.valid = true;
    return GPUMPI_OK;
}

// Reset every piece of shared state the device protocol assumes starts at
// zero, then publish the World.
int init_world(const Config& cfg, void* user_args) {
    WorldAlloc& W = g_world;
    const uint64_t P = (uint64_t)cfg.nranks;
    hipStream_t s = W.stream;
    const bool reset = W.need_reset;
    W.need_reset = true;  // cleared again only when this launch ends cleanly
    if (reset) {
    HIPOK(hipMemsetAsync(W.slots, 0, W.slots_bytes, s));
    HIPOK(hipMemsetAsync(W.chan_head, 0, P * P * sizeof(uint32_t), s));
    HIPOK(hipMemsetAsync(W.done_words, 0, P * (uint64_t)cfg.max_requests * sizeof(uint32_t), s));
    HIPOK(hipMemsetAsync(W.arena_top, 0, P * sizeof(uint64_t), s));
    HIPOK(hipMemsetAsync(W.comm_pool_top, 0, sizeof(uint64_t), s));
    HIPOK(hipMemsetAsync(W.rank_ret, 0, P * sizeof(int), s));
    HIPOK(hipMemsetAsync(W.rank_stats, 0, P * 8 * sizeof(uint64_t), s));
    const int first_ctx = 3;  // 0 null, 1 world, 2 self
    HIPOK(hipMemcpyAsync(W.next_ctx, &first_ctx, sizeof(int), hipMemcpyHostToDevice, s));

    int* world_of = (int*)W.world_tables;
    int* rank_of_world = world_of + P;
    CollSlot* wslots = (CollSlot*)(((uintptr_t)(rank_of_world + P) + 255) & ~(uintptr_t)255);
    CollSlot* sslots = wslots + P;
    std::vector<int> ident(P);
    for (uint64_t i = 0; i < P; i++) ident[i] = (int)i;
    HIPOK(hipMemcpyAsync(world_of, ident.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
    HIPOK(hipMemcpyAsync(rank_of_world, ident.data(), P * sizeof(int), hipMemcpyHostToDevice, s));

    CommShared cw{};
    cw.ctx = 1;
    cw.size = (int)P;
    cw.kind = 0;
    cw.world_of = world_of;
    cw.rank_of_world = rank_of_world;
    cw.slots = wslots;
    HIPOK(hipMemcpyAsync(W.comm_world, &cw, sizeof cw, hipMemcpyHostToDevice, s));
    std::vector<CommShared> cs(P);
    for (uint64_t r = 0; r < P; r++) {
        cs[r] = CommShared{};
        cs[r].ctx = 2;
        cs[r].size = 1;
        cs[r].kind = 1;
        cs[r].slots = sslots + r;
    }
    HIPOK(hipMemcpyAsync(W.comm_self, cs.data(), P * sizeof(CommShared), hipMemcpyHostToDevice, s));
    HIPOK(hipMemsetAsync(W.saves, 0, P * W.save_bytes, s));
    }  // reset

    W.abort_word[0] = 0;
    W.abort_word[1] = 0;  // set by a rank that exits with undelivered messages
    std::memset((void*)W.err, 0, sizeof(ErrorRecord));
    HIPOK(hipMemsetAsync(W.dev_abort, 0, 256, s));
    SvcSlot* svc = nullptr;
    {
        const int rc = gpumpi::host::svc_prepare(cfg.nranks, &svc);
        if (rc) return rc;
    }

    World hw{};
    hw.cfg = cfg;
    hw.slots = W.slots;
    hw.chan_head = W.chan_head;
    hw.done_words = W.done_words;
    hw.staging = W.staging;
    hw.arena = W.arena;
    hw.arena_top = W.arena_top;
    hw.comm_world = W.comm_world;
    hw.comm_self = W.comm_self;
    hw.comm_pool = W.comm_pool;
    hw.comm_pool_top = W.comm_pool_top;
    hw.next_ctx = W.next_ctx;
    hw.rank_ret = W.rank_ret;
    hw.rank_stats = W.rank_stats;
    hw.abort_word = W.abort_word;
    hw.err = W.err;
    hw.user_args = user_args;
    hw.saves = W.saves;
    hw.members_hbm = W.members_hbm;
    hw.dev_abort = W.dev_abort;
    hw.save_bytes = W.save_bytes;
    hw.svc = svc;
    HIPOK(hipMemcpyAsync(W.dworld, &hw, sizeof hw, hipMemcpyHostToDevice, s));
    HIPOK(hipStreamSynchronize(s));
    return GPUMPI_OK;
}

}  // namespace

namespace gpumpi {
namespace host {

int build_config(const gpumpi_config* in, Config* out, int* lds_bytes) {
    const gpumpi_config& c = *in;
    if (c.num_procs < 1) { set_err("num_procs must be >= 1"); return GPUMPI_EINVAL; }
    if (c.threads_per_rank < 64 || c.threads_per_rank > 1024 || c.threads_per_rank % 64) {
        set_err("threads_per_rank must be a multiple of 64 in [64, 1024]");
        return GPUMPI_EINVAL;
    }
    if (c.blocks_per_grid && c.blocks_per_grid != c.num_procs) {
        set_err("one workgroup per rank: blocks_per_grid must equal num_procs");
        return GPUMPI_EINVAL;
    }
    if (!is_pow2(c.chan_slots) || c.chan_slots > kMaxChanSlots) {
        set_err("chan_slots must be a power of two <= %d", kMaxChanSlots);
        return GPUMPI_EINVAL;
    }
    if (c.eager_bytes < 0 || c.eager_bytes > kMaxEagerBytes || c.max_requests < 1 || c.max_requests > 4096) {
        set_err("bad eager_bytes or max_requests");
        return GPUMPI_EINVAL;
    }
    if (c.num_procs > 32767) { set_err("at most 32767 ranks"); return GPUMPI_EINVAL; }
    if (c.comm_pool_bytes < (64u << 10)) { set_err("comm_pool_bytes must be at least 64 KiB"); return GPUMPI_EINVAL; }
    Config& g = *out;
    g = Config{};
    g.nranks = c.num_procs;
    g.threads = c.threads_per_rank;
    g.waves = c.threads_per_rank / 64;
    g.chan_slots = c.chan_slots;
    g.eager_bytes = c.eager_bytes;
    g.slot_bytes = slot_bytes_for(c.eager_bytes);
    g.max_requests = c.max_requests;
    g.helper_min_bytes = c.helper_min_bytes;
    g.staging_bytes = (c.staging_bytes + 511) / 512 * 512;
    g.arena_bytes = ((c.heap_bytes / (size_t)c.num_procs) + 255) / 256 * 256;
    g.comm_pool_bytes = c.comm_pool_bytes;
    g.xcd_map = c.xcd_map;
    g.spin_abort_period = 4096;
    int dev = 0;
    HIPOK(hipGetDevice(&dev));
    int wall_khz = 0, cus = 0;
    HIPOK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
    HIPOK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    g.wall_hz = (uint64_t)wall_khz * 1000ull;
    if (g.wall_hz == 0) g.wall_hz = 100000000ull;
    g.num_xcd = (cus % 8 == 0 && cus >= 64) ? 8 : 1;
    g.user_lds_bytes = c.user_lds_bytes;
    // Collective member tables stay in LDS unless they would keep the ranks
    // from being co-resident: ceil(P / CUs) rank workgroups share a CU's LDS.
    int cu_lds = 0;
    HIPOK(hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    const int per_cu = (g.nranks + cus - 1) / (cus > 0 ? cus : 1);
    const LdsLayout with_members = lds_layout(g.nranks, g.max_requests, g.user_lds_bytes, true);
    g.members_in_lds = (uint64_t)align_up(with_members.total, 1024) * (uint64_t)per_cu <= (uint64_t)cu_lds ? 1 : 0;
    const LdsLayout lay = lds_layout(g.nranks, g.max_requests, g.user_lds_bytes, g.members_in_lds != 0);
    g.lds_bytes = (int)lay.total;
    *lds_bytes = (int)lay.total;
    int max_lds = 0;
    HIPOK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if ((int)lay.total > max_lds) {
        set_err("rank LDS image %u B exceeds the per-workgroup limit %d B (lower max_requests)", lay.total, max_lds);
        return GPUMPI_EINVAL;
    }
    return GPUMPI_OK;
}

}  // namespace host
}  // namespace gpumpi

using gpumpi::host::build_config;

extern "C" {

const char* gpumpi_last_error(void) { return g_err; }

int gpumpi_default_config(gpumpi_config* cfg, int num_procs) {
    std::memset(cfg, 0, sizeof *cfg);
    cfg->num_procs = num_procs;
    cfg->threads_per_rank = 256;
    cfg->blocks_per_grid = 0;
    cfg->max_requests = 256;
    cfg->chan_slots = 8;
    cfg->eager_bytes = 448;
    cfg->helper_min_bytes = 16384;
    cfg->xcd_map = 1;
    cfg->stack_size = 0;
    cfg->heap_bytes = (size_t)num_procs * (4u << 20);
    cfg->staging_bytes = 4u << 20;
    cfg->comm_pool_bytes = 64u << 20;
    cfg->timeout_s = 120.0;
    cfg->verbose = 0;
    return GPUMPI_OK;
}

// Tuning file (written by the autotuner, gpumpi/tune.py): "key value" lines,
// '#' comments; keys threads_per_rank, helper_min_bytes, eager_bytes,
// chan_slots.  Values are checked as build_config checks them.
int gpumpi_apply_tuning(gpumpi_config* cfg, const char* path) {
    if (!cfg || !path) {
        set_err("gpumpi_apply_tuning: null argument");
        return GPUMPI_EINVAL;
    }
    FILE* f = std::fopen(path, "r");
    if (!f) {
        set_err("cannot open tuning file %s", path);
        return GPUMPI_EINVAL;
    }
    gpumpi_config t = *cfg;
    char line[256];
    int lineno = 0, rc = GPUMPI_OK;
    while (rc == GPUMPI_OK && std::fgets(line, sizeof line, f)) {
        lineno++;
        char* h = std::strchr(line, '#');
        if (h) *h = 0;
        char key[64];
        long long v = 0;
        char extra[8];
        const int n = std::sscanf(line, "%63s %lld %7s", key, &v, extra);
        if (n <= 0) continue;  // blank
        if (n != 2) {
            set_err("%s:%d: expected 'key value'", path, lineno);
            rc = GPUMPI_EINVAL;
        } else if (!std::strcmp(key, "threads_per_rank")) {
            if (v < 64 || v > 1024 || v % 64) rc = GPUMPI_EINVAL;
            t.threads_per_rank = (int)v;
        } else if (!std::strcmp(key, "helper_min_bytes")) {
            if (v < 0 || v > (1ll << 30)) rc = GPUMPI_EINVAL;
            t.helper_min_bytes = (int)v;
        } else if (!std::strcmp(key, "eager_bytes")) {
            if (v < 0 || v > kMaxEagerBytes) rc = GPUMPI_EINVAL;
            t.eager_bytes = (int)v;
        } else if (!std::strcmp(key, "chan_slots")) {
            if (v < 1 || v > kMaxChanSlots || !is_pow2((int)v)) rc = GPUMPI_EINVAL;
            t.chan_slots = (int)v;
        } else if (std::strcmp(key, "ranks") && std::strcmp(key, "device") && std::strcmp(key, "objective")) {
            set_err("%s:%d: unknown key %s", path, lineno, key);
            rc = GPUMPI_EINVAL;
            continue;
        }
        if (rc != GPUMPI_OK && n == 2) set_err("%s:%d: bad value %lld for %s", path, lineno, v, key);
    }
    std::fclose(f);
    if (rc == GPUMPI_OK) *cfg = t;
    return rc;
}

int gpumpi_parse_args(int argc, char** argv, gpumpi_config* cfg) {
    int pos = -1;
    for (int i = 0; i < argc; i++)
        if (std::strcmp(argv[i], "---gpumpi") == 0) pos = i;
    if (pos < 0) {
        set_err("gpumpi options must follow '---gpumpi'");
        return -1;
    }
    for (int i = pos + 1; i < argc; i++) {
        const char* a = argv[i];
        auto val = [&](long long* out) -> bool {
            if (i + 1 >= argc) return false;
            char* end = nullptr;
            *out = std::strtoll(argv[++i], &end, 10);
            return end && *end == 0;
        };
        long long v = 0;
        if (!std::strcmp(a, "--tuning")) {
            if (i + 1 >= argc) {
                set_err("bad or missing value for %s", a);
                return -2;
            }
            if (gpumpi_apply_tuning(cfg, argv[++i]) != GPUMPI_OK) return -2;
            continue;
        }
        if (!std::strcmp(a, "-h") || !std::strcmp(a, "--help")) {
            std::printf("GPU MPI options (after ---gpumpi):\n"
                        "  -n N   number of ranks (one workgroup each)\n"
                        "  -b T   threads per rank workgroup (multiple of 64)\n"
                        "  -g G   workgroups (must equal -n)\n"
                        "  -s S   stack bytes per lane\n"
                        "  -p H   device heap bytes (all ranks)\n"
                        "  -e R   pending requests per rank\n"
                        "  -q Q   slots per (src,dst) channel\n"
                        "  -x E   eager message bytes\n"
                        "  -m M   bytes from which a copy is split across the rank's waves\n"
                        "  --tuning FILE  apply an autotuner's choices\n"
                        "  -t S   watchdog seconds (0: none)\n");
            continue;
        }
        bool ok = val(&v);
        if (!ok) {
            set_err("bad or missing value for %s", a);
            return -2;
        }
        if (!std::strcmp(a, "-n") || !std::strcmp(a, "--numProcs")) cfg->num_procs = (int)v;
        else if (!std::strcmp(a, "-b") || !std::strcmp(a, "--threadsPerBlock")) cfg->threads_per_rank = (int)v;
        else if (!std::strcmp(a, "-g") || !std::strcmp(a, "--blocksPerGrid")) cfg->blocks_per_grid = (int)v;
        else if (!std::strcmp(a, "-s") || !std::strcmp(a, "--stackSize")) cfg->stack_size = (size_t)v;
        else if (!std::strcmp(a, "-p") || !std::strcmp(a, "--heapSize")) cfg->heap_bytes = (size_t)v;
        else if (!std::strcmp(a, "-e") || !std::strcmp(a, "--pendingBufferSize")) cfg->max_requests = (int)v;
        else if (!std::strcmp(a, "-q")) cfg->chan_slots = (int)v;
        else if (!std::strcmp(a, "-x")) cfg->eager_bytes = (int)v;
        else if (!std::strcmp(a, "-m")) cfg->helper_min_bytes = (int)v;
        else if (!std::strcmp(a, "-t")) cfg->timeout_s = (double)v;
        else {
            set_err("unknown gpumpi option %s", a);
            return -2;
        }
    }
    if (cfg->blocks_per_grid && !cfg->num_procs) cfg->num_procs = cfg->blocks_per_grid;
    return pos;
}

int gpumpi_max_ranks(const gpumpi_config* cfg, const void* kernel) {
    Config g;
    int lds = 0;
    int rc = build_config(cfg, &g, &lds);
    if (rc) return -rc;
    int per_cu = 0, cus = 0, dev = 0;
    HIPOK(hipGetDevice(&dev));
    HIPOK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPOK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, cfg->threads_per_rank, (size_t)lds));
    return per_cu * cus;
}

}  // extern "C" (reopened below)

namespace {

// Waits for ev1 with the watchdog, serving the ranks' host requests
// (MPI_File_*, printf) meanwhile; returns a GPUMPI status.
int wait_launch(const gpumpi_config* cfg_in, gpumpi_result& R) {
    WorldAlloc& W = g_world;
    const auto t0 = std::chrono::steady_clock::now();
    bool fired = false;
    int busy = 0;  // polls left without sleeping after a served request
    for (;;) {
        if (gpumpi::host::svc_poll() > 0) busy = 2000;
        hipError_t q = hipEventQuery(W.ev1);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) {
            set_err("kernel failed: %s", hipGetErrorString(q));
            return GPUMPI_EHIP;
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!fired && cfg_in->timeout_s > 0 && el > cfg_in->timeout_s) {
            __atomic_store_n(W.abort_word, 1, __ATOMIC_SEQ_CST);
            fired = true;
            R.timed_out = 1;
        }
        if (busy > 0) busy--;
        else std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    float ms = 0.f;
    HIPOK(hipEventElapsedTime(&ms, W.ev0, W.ev1));
    R.kernel_ms = ms;
    return GPUMPI_OK;
}

// Collects return codes and the device error record after a launch.
int finish_launch(const gpumpi_config* cfg_in, const Config& cfg, gpumpi_result& R) {
    WorldAlloc& W = g_world;
    std::vector<int> rets(cfg.nranks);
    HIPOK(hipMemcpy(rets.data(), W.rank_ret, sizeof(int) * cfg.nranks, hipMemcpyDeviceToHost));
    for (int r : rets)
        if (r != 0) R.any_rank_failed = 1;
    const ErrorRecord& e = *W.err;
    R.mpi_error = e.code;
    R.mpi_error_rank = e.rank;
    R.mpi_error_line = e.line;
    R.mpi_error_count = e.count;
    std::memcpy(R.mpi_error_msg, e.msg, sizeof R.mpi_error_msg);
    R.mpi_error_msg[sizeof R.mpi_error_msg - 1] = 0;
    // Channel cursors are reusable only after a clean launch that left no
    // undelivered message behind.
    if (!R.timed_out && !R.mpi_error && __atomic_load_n(W.abort_word + 1, __ATOMIC_SEQ_CST) == 0)
        W.need_reset = false;
    else
        W.need_reset = true;
    if (R.timed_out) {
        set_err("watchdog: launch exceeded %.1f s and was aborted", cfg_in->timeout_s);
        return GPUMPI_ETIMEOUT;
    }
    if (R.mpi_error) {
        set_err("device MPI error %d on rank %d: %s", e.code, e.rank, e.msg);
        return GPUMPI_EDEVICE;
    }
    return GPUMPI_OK;
}

// Everything a launch needs: checked configuration, world, device argv.
struct Prepared {
    Config cfg;
    int lds = 0;
    char* dargv = nullptr;
    World* wptr = nullptr;
    char** argvp = nullptr;
    int argc = 0;
    ~Prepared() {
        if (dargv) (void)hipFree(dargv);
    }
};

int prepare_launch(const gpumpi_config* cfg_in, const void* kernel, int argc, char** argv, void* user_args,
                   gpumpi_result& R, Prepared& L) {
    // A GPU error left by earlier work (another launch, a copy, a fill) is
    // reported as such, before anything of this launch runs.
    {
        const hipError_t pend = hipDeviceSynchronize();
        if (pend != hipSuccess) {
            set_err("GPU error pending before this launch: %s", hipGetErrorString(pend));
            return GPUMPI_EHIP;
        }
    }
    int rc = build_config(cfg_in, &L.cfg, &L.lds);
    if (rc) return rc;
    R.lds_bytes = L.lds;
    if (L.lds > 48 * 1024) HIPOK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, L.lds));
    const int maxr = gpumpi_max_ranks(cfg_in, kernel);
    if (maxr < L.cfg.nranks) {
        set_err("%d ranks of %d threads cannot all be resident (max %d): grid barriers would deadlock",
                L.cfg.nranks, L.cfg.threads, maxr);
        return GPUMPI_ETOOBIG;
    }
    if ((rc = alloc_world(L.cfg))) return rc;
    if ((rc = init_world(L.cfg, user_args))) return rc;
    if (cfg_in->stack_size) HIPOK(hipDeviceSetLimit(hipLimitStackSize, cfg_in->stack_size));
    // argv in device memory (main.cu:23-42 used managed memory).
    size_t strbytes = 0;
    for (int i = 0; i < argc; i++) strbytes += std::strlen(argv[i]) + 1;
    const size_t ptrbytes = (size_t)(argc + 1) * sizeof(char*);
    std::vector<uint8_t> img(ptrbytes + strbytes + 1);
    HIPOK(hipMalloc(&L.dargv, img.size()));
    char** ptrs = (char**)img.data();
    size_t off = ptrbytes;
    for (int i = 0; i < argc; i++) {
        ptrs[i] = L.dargv + off;
        std::memcpy(img.data() + off, argv[i], std::strlen(argv[i]) + 1);
        off += std::strlen(argv[i]) + 1;
    }
    ptrs[argc] = nullptr;
    HIPOK(hipMemcpy(L.dargv, img.data(), img.size(), hipMemcpyHostToDevice));
    L.wptr = g_world.dworld;
    L.argvp = (char**)L.dargv;
    L.argc = argc;
    return GPUMPI_OK;
}

}  // namespace

extern "C" {

int gpumpi_run(const gpumpi_config* cfg_in, const void* kernel, int argc, char** argv, void* user_args,
               gpumpi_result* res) {
    gpumpi_result local{};
    gpumpi_result& R = res ? *res : local;
    std::memset(&R, 0, sizeof R);
    Prepared L;
    int rc = prepare_launch(cfg_in, kernel, argc, argv, user_args, R, L);
    if (rc) return R.status = rc;
    WorldAlloc& W = g_world;
    void* params[] = {(void*)&L.wptr, (void*)&L.argc, (void*)&L.argvp};
    HIPOK(hipEventRecord(W.ev0, W.stream));
    hipError_t le = hipLaunchCooperativeKernel(kernel, dim3((unsigned)L.cfg.nranks), dim3((unsigned)L.cfg.threads),
                                               params, (unsigned)L.lds, W.stream);
    if (le != hipSuccess) {
        (void)hipGetLastError();
        set_err("cooperative launch of %d x %d failed: %s", L.cfg.nranks, L.cfg.threads, hipGetErrorString(le));
        return R.status = (le == hipErrorCooperativeLaunchTooLarge) ? GPUMPI_ETOOBIG : GPUMPI_EHIP;
    }
    HIPOK(hipEventRecord(W.ev1, W.stream));
    if ((rc = wait_launch(cfg_in, R))) return R.status = rc;
    return R.status = finish_launch(cfg_in, L.cfg, R);
}

int gpumpi_run_graph(const gpumpi_config* cfg_in, const void* kernel, int argc, char** argv, void* user_args,
                     int replays, gpumpi_result* res, double* replay_ms) {
    gpumpi_result local{};
    gpumpi_result& R = res ? *res : local;
    std::memset(&R, 0, sizeof R);
    if (replays < 1) {
        set_err("replays must be >= 1");
        return R.status = GPUMPI_EINVAL;
    }
    Prepared L;
    int rc = prepare_launch(cfg_in, kernel, argc, argv, user_args, R, L);
    if (rc) return R.status = rc;
    WorldAlloc& W = g_world;
    // Capture one launch of the rank kernel (residency was checked above, as
    // the cooperative launch would) and replay it.
    void* params[] = {(void*)&L.wptr, (void*)&L.argc, (void*)&L.argvp};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPOK(hipStreamBeginCapture(W.stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipLaunchKernel(kernel, dim3((unsigned)L.cfg.nranks), dim3((unsigned)L.cfg.threads), params,
                                    (size_t)L.lds, W.stream);
    hipError_t ce = hipStreamEndCapture(W.stream, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
        set_err("graph capture failed: %s / %s", hipGetErrorString(le), hipGetErrorString(ce));
        if (graph) (void)hipGraphDestroy(graph);
        return R.status = GPUMPI_EHIP;
    }
    HIPOK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    double total = 0;
    for (int i = 0; i < replays && rc == GPUMPI_OK; i++) {
        W.abort_word[0] = 0;
        W.abort_word[1] = 0;
        std::memset((void*)W.err, 0, sizeof(ErrorRecord));
        // Between replays, outside the captured graph.
        HIPOK(hipMemsetAsync(W.dev_abort, 0, 256, W.stream));
        HIPOK(hipEventRecord(W.ev0, W.stream));
        HIPOK(hipGraphLaunch(exec, W.stream));
        HIPOK(hipEventRecord(W.ev1, W.stream));
        if ((rc = wait_launch(cfg_in, R))) break;
        rc = finish_launch(cfg_in, L.cfg, R);
        if (replay_ms) replay_ms[i] = R.kernel_ms;
        total += R.kernel_ms;
        if (W.need_reset && i + 1 < replays && rc == GPUMPI_OK) {
            set_err("a replay left undelivered messages; later replays would see them");
            rc = GPUMPI_EDEVICE;
        }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    R.kernel_ms = total / replays;
    return R.status = rc;
}

int gpumpi_main(int argc, char** argv, const void* kernel) {
    gpumpi_config cfg;
    gpumpi_default_config(&cfg, 0);
    int user_argc = gpumpi_parse_args(argc, argv, &cfg);
    if (user_argc < 0) {
        std::fprintf(stderr, "GPUMPI: %s\n", gpumpi_last_error());
        return 1;
    }
    if (cfg.num_procs <= 0) cfg.num_procs = 1;
    if (cfg.heap_bytes == 0) cfg.heap_bytes = (size_t)cfg.num_procs * (4u << 20);
    gpumpi_result r;
    int rc = gpumpi_run(&cfg, kernel, user_argc, argv, nullptr, &r);
    if (rc != GPUMPI_OK) {
        std::fprintf(stderr, "GPUMPI: %s\n", gpumpi_last_error());
        return 1;
    }
    return r.any_rank_failed ? 1 : 0;
}

int gpumpi_reset_world(void) {
    g_world.need_reset = true;
    return 0;
}

int gpumpi_release_world(void) {
    if (g_world.valid) {
        (void)hipDeviceSynchronize();
        g_world.release();
    }
    return 0;
}

int gpumpi_device_alloc(void** ptr, size_t bytes) {
    HIPOK(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}
int gpumpi_device_free(void* ptr) {
    HIPOK(hipFree(ptr));
    return 0;
}
int gpumpi_memcpy_htod(void* dst, const void* src, size_t bytes) {
    HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int gpumpi_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int gpumpi_memset(void* dst, int value, size_t bytes) {
    HIPOK(hipMemset(dst, value, bytes));
    return 0;
}

}  // extern "C"
