// Fiber scheduler and host API of the CPU simulation (hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <sanitizer/asan_interface.h>
#include <sys/mman.h>

#include <chrono>
#include <map>
#include <vector>

// Dynamic LDS of the running workgroup (device.hpp declares the symbol as
// the extern LDS array; the shim turns it into this pointer).
namespace gpumpi {
namespace rt {
__attribute__((aligned(16))) uint8_t (*gpumpi_host_lds)[] = nullptr;
}
}  // namespace gpumpi

namespace simt {

struct Wave {
    uint64_t vals[64];
    uint64_t res[64];
    int args[64];
    const void* site[64] = {};   // call site of each lane's latest wave operation
    int op = 0;
    uint64_t nop = 0;            // sequence number (per lane) of the pending operation
    int arrived = 0;             // lanes waiting at the pending wave operation
    int stepping = 0;            // lanes waiting at the memory-step barrier
    int done = 0;                // lanes that finished the kernel
    int parked = 0;              // lanes waiting at the workgroup barrier
    uint64_t step_gen = 0;
    uint64_t step_since = 0;     // scheduler pass when the pending step started
    uint64_t gen = 0;
    uint64_t open_since = 0;     // scheduler pass when the pending operation started
};
struct Block {
    std::vector<Wave> waves;
    int arrived = 0;
    uint64_t gen = 0;
    int nthreads = 0;
    uint8_t* lds = nullptr;
    size_t lds_bytes = 0;
};

static void wave_check(Wave& w);
Fiber main_fiber;
Fiber* cur = &main_fiber;
dim3 grid_dim, block_dim;
static Fiber* sched_fiber = &main_fiber;
static uint64_t g_counter = 0;
static const void* g_kernel = nullptr;
static void** g_args = nullptr;
static size_t g_stack_bytes = 256 << 10;
static uint64_t g_switches = 0;
static uint64_t g_pass = 0;

[[noreturn]] void fail(const char* what) {
    std::fprintf(stderr, "simt: %s (block %u thread %u)\n", what, cur->bid.x, cur->tid.x);
    std::fflush(stderr);
    std::abort();
}

uint64_t steady_counter() { return ++g_counter; }

// Callee-saved register switch (System V x86-64): saves rbx, rbp, r12-r15 on
// the current stack, stores rsp in *from, loads rsp from to, restores.
extern "C" void simt_switch(void** from, void* to);
asm(R"(
.text
.globl simt_switch
.type simt_switch,@function
simt_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size simt_switch,.-simt_switch
)");

static void switch_to(Fiber* to, bool dying) {
    Fiber* from = cur;
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(dying ? nullptr : &fake, to->stack_lo, to->stack_size);
    cur = to;
    g_switches++;
    simt_switch(&from->sp, to->sp);
    // back on `from`'s stack
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
    cur = from;
}

void yield() {
    if (cur == sched_fiber) return;
    switch_to(sched_fiber, false);
}

// Blocks the current fiber until *gen changes (the scheduler does not
// resume it before then).
static void wait_change(const uint64_t* gen, uint64_t val) {
    while (*gen == val) {
        cur->wait_gen = gen;
        cur->wait_val = val;
        yield();
        cur->wait_gen = nullptr;
    }
}

static void run_kernel_body() {
    typedef void (*RankKernel)(void*, int, char**);
    ((RankKernel)g_kernel)(*(void**)g_args[0], *(int*)g_args[1], *(char***)g_args[2]);
}

extern "C" void simt_fiber_entry() {
    __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
    run_kernel_body();
    cur->done = true;
    cur->wave->done++;
    wave_check(*cur->wave);
    switch_to(sched_fiber, true);
    fail("finished fiber resumed");
}

// Completes the pending wave operation once every live lane has arrived at
// it, or releases the lanes waiting at the memory-step barrier once every
// live lane is at that barrier or at the wave operation.
static void wave_check(Wave& w) {
    const int live = 64 - w.done;
    if (w.arrived > 0 && w.arrived == live) {
        uint64_t ballot = 0;
        if (w.op == OP_BALLOT)
            for (int l = 0; l < 64; l++) ballot |= (w.vals[l] ? 1ull : 0ull) << l;
        for (int l = 0; l < 64; l++) {
            switch (w.op) {
            case OP_FIRST: w.res[l] = w.vals[0]; break;
            case OP_BALLOT: w.res[l] = ballot; break;
            case OP_SHFL: w.res[l] = w.vals[(unsigned)w.args[l] & 63]; break;
            case OP_XOR: w.res[l] = w.vals[(l ^ w.args[l]) & 63]; break;
            case OP_UP: w.res[l] = l >= w.args[l] ? w.vals[l - w.args[l]] : w.vals[l]; break;
            default: w.res[l] = 0; break;
            }
        }
        w.arrived = 0;
        w.gen++;
    } else if (w.stepping > 0 && w.stepping + w.arrived + w.parked == live) {
        w.stepping = 0;
        w.step_gen++;
    }
}

// Lockstep at memory-operation granularity: every load and store of the
// device code (sanitizer-coverage callbacks) waits until all live lanes of
// the wave are at their next memory operation (or at a wave operation), so
// a wave's lanes perform their k-th access before any performs its k+1-th,
// as in the hardware's lockstep execution.
// Device allocations of at least kBulkBytes hold payloads, slots, heaps and
// staging: lanes touch disjoint parts of them between wave-wide handoffs,
// so accesses there (and to a lane's own stack) skip the lockstep step.
// Control state -- LDS, and the small allocations (counters, cursors, done
// words, tables) -- keeps it.
static constexpr size_t kBulkBytes = 8u << 20;
static std::map<uintptr_t, size_t> g_bulk;
static bool needs_step(const void* p) {
    const uintptr_t a = (uintptr_t)p;
    Fiber* f = cur;
    if (f->block && a >= (uintptr_t)f->block->lds && a < (uintptr_t)f->block->lds + f->block->lds_bytes) return true;
    if (a >= (uintptr_t)f->stack_lo && a < (uintptr_t)f->stack_lo + f->stack_size) return false;
    auto it = g_bulk.upper_bound(a);
    if (it != g_bulk.begin()) {
        --it;
        if (a < it->first + it->second) return false;
    }
    return true;
}
void note_alloc(void* p, size_t n) {
    if (n >= kBulkBytes) g_bulk[(uintptr_t)p] = n;
}
void note_free(void* p) { g_bulk.erase((uintptr_t)p); }

void mem_step(const void* addr) {
    Fiber* f = cur;
    if (!f->wave) return;
    if (!needs_step(addr)) return;
    Wave& w = *f->wave;
    const uint64_t g = w.step_gen;
    if (w.stepping == 0) w.step_since = g_pass;
    w.stepping++;
    wave_check(w);
    wait_change(&w.step_gen, g);
}

__attribute__((noinline)) uint64_t wave_op(int op, uint64_t v, int arg) {
    Fiber* f = cur;
    if (!f->wave) fail("wave operation outside a rank kernel");
    Wave& w = *f->wave;
    w.site[f->lane] = __builtin_return_address(0);
    f->nops++;
    if (w.arrived == 0) {
        w.op = op;
        w.nop = f->nops;
        w.open_since = g_pass;
    } else if (w.op != op || w.nop != f->nops) {
        int first = -1;
        for (int l = 0; l < 64; l++)
            if (l != f->lane && w.site[l]) { first = l; break; }
        std::fprintf(stderr, "simt: lane %d reached wave operation #%llu (kind %d) at %p while the wave waits at #%llu "
                     "(kind %d, lane %d at %p): lanes took different numbers of wave operations\n",
                     f->lane, (unsigned long long)f->nops, op, w.site[f->lane], (unsigned long long)w.nop, w.op, first,
                     first >= 0 ? w.site[first] : nullptr);
        __sanitizer_print_stack_trace();
        fail("wave divergence at a wave operation");
    }
    w.vals[f->lane] = v;
    w.args[f->lane] = arg;
    const uint64_t g = w.gen;
    w.arrived++;
    wave_check(w);
    wait_change(&w.gen, g);
    return w.res[f->lane];
}

void block_barrier() {
    Fiber* f = cur;
    if (!f->block) return;  // plain kernels: one thread at a time
    Block& b = *f->block;
    const uint64_t g = b.gen;
    if (++b.arrived == b.nthreads) {
        b.arrived = 0;
        b.gen++;
    } else {
        // Parked lanes do not hold up their wave's memory steps.
        f->wave->parked++;
        wave_check(*f->wave);
        wait_change(&b.gen, g);
        f->wave->parked--;
    }
}

bool is_private(const void* p) {
    const char* c = (const char*)p;
    return cur->stack_lo && c >= cur->stack_lo && c < cur->stack_lo + cur->stack_size;
}
bool is_shared(const void* p) {
    if (!cur->block) return false;
    const uint8_t* c = (const uint8_t*)p;
    return c >= cur->block->lds && c < cur->block->lds + cur->block->lds_bytes;
}

static Fiber plain;
void plain_begin(dim3 grid, dim3 block) {
    grid_dim = grid;
    block_dim = block;
    cur = &plain;
}
void plain_thread(unsigned bx, unsigned by, unsigned t) {
    plain.bid = dim3(bx, by, 0);
    plain.tid = dim3(t, 0, 0);
    plain.lane = (int)(t & 63);
}
void plain_end() { cur = &main_fiber; }

static int launch(const void* k, dim3 grid, dim3 block, void** args, size_t lds) {
    if (block.x % 64) return hipErrorInvalidValue;
    grid_dim = grid;
    block_dim = block;
    g_kernel = k;
    g_args = args;
    const unsigned nb = grid.x, nt = block.x;
    std::vector<Block> blocks(nb);
    std::vector<Fiber> fibers((size_t)nb * nt);
    const size_t S = g_stack_bytes;
    char* stacks = (char*)mmap(nullptr, S * fibers.size(), PROT_READ | PROT_WRITE,
                               MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (stacks == MAP_FAILED) return hipErrorOutOfMemory;
    for (unsigned b = 0; b < nb; b++) {
        Block& B = blocks[b];
        B.waves.resize(nt / 64);
        B.nthreads = (int)nt;
        B.lds_bytes = lds;
        B.lds = (uint8_t*)std::malloc(lds ? lds : 16);
        std::memset(B.lds, 0xCD, lds);
        for (unsigned t = 0; t < nt; t++) {
            Fiber& f = fibers[(size_t)b * nt + t];
            f.bid = dim3(b, 0, 0);
            f.tid = dim3(t, 0, 0);
            f.lane = (int)(t & 63);
            f.wave = &B.waves[t / 64];
            f.block = &B;
            f.stack_lo = stacks + S * ((size_t)b * nt + t);
            f.stack_size = S;
            // Initial frame: six callee-saved registers, then the entry as
            // return address; the entry starts with rsp = 8 mod 16.
            uint64_t* top = (uint64_t*)(f.stack_lo + S - 64);
            for (int i = 0; i < 6; i++) top[i] = 0;
            top[6] = (uint64_t)(uintptr_t)&simt_fiber_entry;
            f.sp = top;
        }
    }
    size_t live = fibers.size();
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t passes = 0;
    const char* tl = std::getenv("SIMT_TIMEOUT");
    const double limit = tl ? std::atof(tl) : 1800.0;
    while (live) {
        live = 0;
        for (Fiber& f : fibers) {
            if (f.done) continue;
            if (f.wait_gen && *f.wait_gen == f.wait_val) {
                live++;
                continue;
            }
            gpumpi::rt::gpumpi_host_lds = (uint8_t(*)[])f.block->lds;
            switch_to(&f, false);
            if (!f.done) live++;
        }
        passes++;
        g_pass = passes;
        if ((passes & 255) == 0) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            bool stuck = el > limit;
            for (Block& B : blocks)
                for (Wave& W : B.waves)
                    if ((W.arrived > 0 && passes - W.open_since > 20000) ||
                        (W.stepping > 0 && passes - W.step_since > 20000))
                        stuck = true;
            if (stuck) {
                std::fprintf(stderr, "simt: no progress (%.0f s, %llu passes); waves with a pending operation:\n", el,
                             (unsigned long long)passes);
                for (unsigned b = 0; b < nb; b++)
                    for (size_t wi = 0; wi < blocks[b].waves.size(); wi++) {
                        Wave& W = blocks[b].waves[wi];
                        if (W.arrived == 0 && W.stepping == 0) continue;
                        std::fprintf(stderr, "  block %u wave %zu op %d arrived %d stepping %d parked %d done %d since pass %llu/%llu\n", b, wi, W.op,
                                     W.arrived, W.stepping, W.parked, W.done, (unsigned long long)W.open_since,
                                     (unsigned long long)W.step_since);
                        const void* last = nullptr;
                        for (int l = 0; l < 64; l++)
                            if (W.site[l] != last) {
                                std::fprintf(stderr, "    lane %d site %p\n", l, W.site[l]);
                                last = W.site[l];
                            }
                    }
                fail("simulated launch stuck");
            }
        }
    }
    munmap(stacks, S * fibers.size());
    for (Block& B : blocks) std::free(B.lds);
    cur = &main_fiber;
    return hipSuccess;
}

}  // namespace simt

extern "C" {
#define SIMT_STEP(n, T) \
    void __sanitizer_cov_load##n(T* p) { simt::mem_step(p); } \
    void __sanitizer_cov_store##n(T* p) { simt::mem_step(p); }
SIMT_STEP(1, uint8_t)
SIMT_STEP(2, uint16_t)
SIMT_STEP(4, uint32_t)
SIMT_STEP(8, uint64_t)
SIMT_STEP(16, __int128)
#undef SIMT_STEP
}

// ---------------------------------------------------------------------------
// Host API
// ---------------------------------------------------------------------------
struct simt_stream { int x; };
struct simt_event { double t; };
static simt_stream g_stream;
static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

hipError_t hipMalloc(void** p, size_t n) {
    void* q = nullptr;
    if (posix_memalign(&q, 256, n ? n : 1)) return hipErrorOutOfMemory;
    std::memset(q, 0xA5, n < (64u << 20) ? n : (64u << 20));  // device memory is not zeroed
    *p = q;
    simt::note_alloc(q, n);
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    simt::note_free(p);
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipMemset(void* p, int v, size_t n) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { std::memset(p, v, n); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = &g_stream; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) { *e = new simt_event{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = now_ms(); return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(b->t - a->t); return hipSuccess; }
hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    switch (a) {
    case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
    case hipDeviceAttributeWallClockRate: *v = 100000; break;
    default: *v = 160 * 1024; break;
    }
    return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, const void*, int threads, size_t lds) {
    int by_lds = lds ? (int)((160 * 1024) / lds) : 32;
    int by_waves = 8 / ((threads + 63) / 64 > 0 ? (threads + 63) / 64 : 1);  // 184 VGPRs: 2 waves per SIMD
    *n = by_lds < by_waves ? by_lds : by_waves;
    if (*n < 1) *n = 0;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipDeviceSetLimit(hipLimit_t, size_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "simulated HIP error"; }
hipError_t hipLaunchCooperativeKernel(const void* k, dim3 grid, dim3 block, void** args, unsigned lds, hipStream_t) {
    return simt::launch(k, grid, block, args, lds);
}
hipError_t hipLaunchKernel(const void* k, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t) {
    return simt::launch(k, grid, block, args, lds);
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
