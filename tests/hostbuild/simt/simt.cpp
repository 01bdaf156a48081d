// Fiber scheduler and host API of the CPU simulation (hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <sanitizer/asan_interface.h>
#include <sys/mman.h>

#include <sanitizer/common_interface_defs.h>

#include <chrono>
#include <map>
#include <unordered_map>
#include <unordered_set>
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

[[noreturn]] static void finish_lane() {
    cur->done = true;
    cur->wave->done++;
    wave_check(*cur->wave);
    switch_to(sched_fiber, true);
    fail("finished fiber resumed");
}

extern "C" void simt_fiber_entry() {
    __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
    run_kernel_body();
    finish_lane();
}

// s_endpgm: the lane's fiber ends without returning through its frames.
void lane_exit() { finish_lane(); }

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

// ---------------------------------------------------------------------------
// Happens-before checker (SIMT_RACE=1).
//
// Threads are ranks (workgroups): the waves of one rank synchronise through
// LDS and are not checked against each other.  Each rank carries a vector
// clock.  Synchronisation follows the scoped-atomic model of the device
// code, agent or system scope only (a workgroup-scope fence or atomic
// orders nothing between ranks):
//   release fence        rel[A] = VC[A]; VC[A][A]++
//   atomic store/RMW     sync[x] = rel[A] (store) or sync[x] |= rel[A] (RMW),
//                        after a release-fence step of its own if the
//                        operation is itself release
//   atomic load/RMW      acquire: VC[B] |= sync[x]; relaxed: pend[B] |= sync[x]
//   acquire fence        VC[B] |= pend[B]
// Every plain load and store of global memory (4-byte words; 1- and 2-byte
// accesses by byte) keeps the last write epoch and the read epochs since
// then; an access that conflicts with another rank's access it is not
// ordered after is reported with both sites.  A launch starts from a clean
// state: the host's writes before it, and everything of the launch before,
// are ordered by the launch itself.
// ---------------------------------------------------------------------------
static bool g_race = false;
static int g_nranks = 0;
static std::vector<std::vector<uint32_t>> g_vc, g_rel, g_pend;
struct RaceEpoch {
    int32_t r = -1;
    uint32_t c = 0;
    const void* pc = nullptr;
};
struct RaceShadow {
    RaceEpoch w;
    RaceEpoch rd;                       // the one read epoch, or
    std::vector<RaceEpoch>* rv = nullptr;  // per-rank read epochs once reads are concurrent
};
static std::unordered_map<uintptr_t, RaceShadow> g_shadow;
static std::unordered_map<uintptr_t, std::vector<uint32_t>> g_sync;
static std::unordered_set<uint64_t> g_race_seen;
static uint64_t g_race_reports = 0;

static void race_reset(int nranks) {
    const char* e = std::getenv("SIMT_RACE");
    g_race = e && e[0] == '1';
    if (!g_race) return;
    g_nranks = nranks;
    g_vc.assign(nranks, std::vector<uint32_t>(nranks, 0));
    g_rel = g_vc;
    g_pend = g_vc;
    for (int r = 0; r < nranks; r++) g_vc[r][r] = 1;
    for (auto& kv : g_shadow) delete kv.second.rv;
    g_shadow.clear();
    g_sync.clear();
}
static void join(std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    for (size_t i = 0; i < a.size() && i < b.size(); i++)
        if (b[i] > a[i]) a[i] = b[i];
}
static int race_rank() {
    if (!g_race || !cur->block || cur->in_atomic) return -1;
    return (int)cur->bid.x;
}
static void race_report(const char* what, uintptr_t key, const RaceEpoch& prev, int me, const void* pc) {
    const uint64_t sig = (uint64_t)(uintptr_t)prev.pc * 1000003u ^ (uint64_t)(uintptr_t)pc;
    g_race_reports++;
    if (!g_race_seen.insert(sig).second || g_race_seen.size() > 40) return;
    char a[512], b[512];
    __sanitizer_symbolize_pc(const_cast<void*>(pc), "%F %L", a, sizeof a);
    __sanitizer_symbolize_pc(const_cast<void*>(prev.pc), "%F %L", b, sizeof b);
    std::fprintf(stderr,
                 "simt race: %s at %#lx: rank %d (clock of rank %d: %u) is not ordered after rank %d's access "
                 "(epoch %u)\n  this access: %s\n  other access: %s\n",
                 what, (unsigned long)(key & ~(1ull << 63)), me, prev.r, g_vc[me][prev.r], prev.r, prev.c, a, b);
}
static bool race_checked(const void* p) {
    Fiber* f = cur;
    const uintptr_t a = (uintptr_t)p;
    if (a >= (uintptr_t)f->block->lds && a < (uintptr_t)f->block->lds + f->block->lds_bytes) return false;
    if (a >= (uintptr_t)f->stack_lo && a < (uintptr_t)f->stack_lo + f->stack_size) return false;
    return true;
}
static void race_word(uintptr_t key, bool store, int me, const void* pc) {
    RaceShadow& sh = g_shadow[key];
    const std::vector<uint32_t>& vc = g_vc[me];
    if (sh.w.r >= 0 && sh.w.r != me && sh.w.c > vc[sh.w.r])
        race_report(store ? "write after unordered write" : "read of an unordered write", key, sh.w, me, pc);
    if (store) {
        if (sh.rv) {
            for (const RaceEpoch& e : *sh.rv)
                if (e.r >= 0 && e.r != me && e.c > vc[e.r]) {
                    race_report("write after an unordered read", key, e, me, pc);
                    break;
                }
            delete sh.rv;
            sh.rv = nullptr;
        } else if (sh.rd.r >= 0 && sh.rd.r != me && sh.rd.c > vc[sh.rd.r]) {
            race_report("write after an unordered read", key, sh.rd, me, pc);
        }
        sh.rd = RaceEpoch();
        sh.w = RaceEpoch{me, vc[me], pc};
        return;
    }
    const RaceEpoch now{me, vc[me], pc};
    if (sh.rv) {
        if ((size_t)me < sh.rv->size()) (*sh.rv)[me] = now;
    } else if (sh.rd.r < 0 || sh.rd.r == me || sh.rd.c <= vc[sh.rd.r]) {
        sh.rd = now;
    } else {
        sh.rv = new std::vector<RaceEpoch>(g_nranks);
        (*sh.rv)[sh.rd.r] = sh.rd;
        (*sh.rv)[me] = now;
    }
}
static void race_access(const void* p, int n, bool store, const void* pc) {
    const int me = race_rank();
    if (me < 0 || !race_checked(p)) return;
    const uintptr_t a = (uintptr_t)p;
    if (n < 4) {
        for (int i = 0; i < n; i++) race_word((a + i) | (1ull << 63), store, me, pc);
        return;
    }
    for (uintptr_t w = a & ~(uintptr_t)3; w < a + n; w += 4) race_word(w, store, me, pc);
}
static bool is_release(int o) { return o == __ATOMIC_RELEASE || o == __ATOMIC_ACQ_REL || o == __ATOMIC_SEQ_CST; }
static bool is_acquire(int o) {
    return o == __ATOMIC_ACQUIRE || o == __ATOMIC_ACQ_REL || o == __ATOMIC_SEQ_CST || o == __ATOMIC_CONSUME;
}
static void release_step(int me) {
    g_rel[me] = g_vc[me];
    g_vc[me][me]++;
}
void race_fence(int order, const char* scope) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (!g_race || !cur->block) return;
    if (std::strcmp(scope, "agent") != 0 && scope[0] != 0) return;  // workgroup / wavefront
    const int me = (int)cur->bid.x;
    if (is_release(order)) release_step(me);
    if (is_acquire(order)) join(g_vc[me], g_pend[me]);
}
void race_atomic(const void* p, int kind, int order, int scope) {
    if (!g_race || !cur->block || scope < 4) return;  // below agent scope: rank-local
    const int me = (int)cur->bid.x;
    const uintptr_t key = (uintptr_t)p;
    if (kind != RACE_STORE) {
        auto it = g_sync.find(key);
        if (it != g_sync.end()) join(is_acquire(order) ? g_vc[me] : g_pend[me], it->second);
    }
    if (kind != RACE_LOAD) {
        if (is_release(order)) release_step(me);
        std::vector<uint32_t>& sy = g_sync[key];
        if (kind == RACE_STORE || sy.empty()) sy = g_rel[me];
        else join(sy, g_rel[me]);
    }
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

// Private (scratch) addresses as the GPU has them: one address names every
// lane's own copy.  Here each lane's stack is its own region, all of one size
// and with the same frames at the same offsets (the lanes run one program),
// so an address in any lane's stack is private, and the calling lane's copy
// of it lies at the same offset in its own stack.
static char* g_stacks = nullptr;
static char* g_stacks_end = nullptr;
bool is_private(const void* p) {
    const char* c = (const char*)p;
    if (cur->stack_lo && c >= cur->stack_lo && c < cur->stack_lo + cur->stack_size) return true;
    return g_stacks && c >= g_stacks && c < g_stacks_end;
}
void* lane_private(const void* p) {
    const char* c = (const char*)p;
    if (!cur->stack_lo || !g_stacks || c < g_stacks || c >= g_stacks_end) return (void*)p;
    return cur->stack_lo + (size_t)(c - g_stacks) % cur->stack_size;
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

// The launcher's service-channel poll (host/service.cpp), when linked: the
// simulated launch runs inside hipLaunchCooperativeKernel, so the scheduler
// serves the ranks' host requests itself, as the launcher's wait loop does
// on the GPU.
extern "C" int gpumpi_service_poll(void) __attribute__((weak));

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
    g_stacks = stacks;
    g_stacks_end = stacks + S * fibers.size();
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
    race_reset((int)nb);
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
        if (gpumpi_service_poll && (passes & 15) == 0) {
            Fiber* keep = cur;
            cur = &main_fiber;
            gpumpi_service_poll();
            cur = keep;
        }
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
    g_stacks = g_stacks_end = nullptr;
    munmap(stacks, S * fibers.size());
    for (Block& B : blocks) std::free(B.lds);
    cur = &main_fiber;
    if (g_race) {
        std::fprintf(stderr, "simt race check: %llu report(s) in this launch (%zu distinct site pairs)\n",
                     (unsigned long long)g_race_reports, g_race_seen.size());
        g_race_reports = 0;
        g_race_seen.clear();
    }
    return hipSuccess;
}

}  // namespace simt

extern "C" {
#define SIMT_STEP(n, T)                                                              \
    void __sanitizer_cov_load##n(T* p) {                                             \
        simt::mem_step(p);                                                           \
        if (simt::g_race) simt::race_access(p, n, false, __builtin_return_address(0)); \
    }                                                                                \
    void __sanitizer_cov_store##n(T* p) {                                            \
        simt::mem_step(p);                                                           \
        if (simt::g_race) simt::race_access(p, n, true, __builtin_return_address(0));  \
    }
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
