// World state shared by the host launcher and the device runtime.
//
// Everything here lives in HBM (one allocation set per launch shape) except
// the abort word and the error record, which are pinned host memory so a
// host watchdog can stop a running kernel and read why it stopped.
//
// Replaces the reference's managed-memory SharedState
// (gpu_libs/gpu_libc/cuda_mpi.cuh:459-533) and per-thread
// ThreadPrivateState (cuda_mpi.cuh:254-370): rank-private state moves into
// the rank workgroup's LDS (see lds.hpp), rank-to-rank state becomes
// per-(src,dst) channels below.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace gpumpi {
namespace rt {

constexpr int kMaxWaves = 16;        // 1024-thread workgroup
constexpr int kMaxChanSlots = 64;    // consumed-mask width per channel
constexpr int kMaxCommSlots = 8;     // words of per-member collective slot

// Launch-time configuration (copied into the World in HBM).
struct Config {
    int nranks;           // MPI world size = workgroups
    int threads;          // threads per rank workgroup (multiple of 64)
    int waves;            // threads / 64
    int chan_slots;       // slots per (src,dst) channel, power of two <= 64
    int slot_bytes;       // slot_words_for(eager_bytes) * 8, a multiple of 64
    int eager_bytes;      // largest message carried inside a slot
    int max_requests;     // request pool per rank (LDS)
    int helper_min_bytes; // copies at least this large use every wave
    uint64_t staging_bytes; // per-rank bounce buffer for private/LDS buffers
    uint64_t arena_bytes;   // per-rank device heap (gpumpi_malloc)
    uint64_t wall_hz;       // s_memrealtime frequency
    int xcd_map;          // 1: consecutive ranks share an XCD
    int num_xcd;          // XCDs the dispatcher round-robins over
    uint64_t comm_pool_bytes;
    int lds_bytes;        // dynamic LDS per workgroup
    int spin_abort_period; // polls between abort-word checks
    int user_lds_bytes;    // LDS reserved for the rank program (gpumpi_user_lds)
    int members_in_lds;    // 1: collective member tables in LDS; 0: in World::members_hbm
};

// Envelope word of a slot: [63:32] sequence+1, [31:16] context, [15:0] tag.
// Written last, with release semantics, by the sender.
__host__ __device__ inline uint64_t pack_env(uint32_t seq1, int ctx, int tag) {
    return ((uint64_t)seq1 << 32) | ((uint64_t)(uint16_t)ctx << 16) | (uint64_t)(uint16_t)tag;
}
__host__ __device__ inline uint32_t env_seq(uint64_t e) { return (uint32_t)(e >> 32); }
__host__ __device__ inline int env_ctx(uint64_t e) { return (int)(int16_t)(uint16_t)(e >> 16); }
__host__ __device__ inline int env_tag(uint64_t e) { return (int)(int16_t)(uint16_t)e; }

enum SlotFlags : uint32_t { SLOT_EAGER = 1u, SLOT_RNDV = 2u };

// A channel slot is an array of 8-byte words, each stored and loaded whole
// (single-copy atomic) and each carrying the message's sequence number+1 in
// its upper half: [63:32] seq+1, [31:0] value.  A receiver accepts a word
// only when its upper half is the sequence it expects, so a slot needs no
// fences between payload and envelope and no ordering among its words: the
// sender writes every word at once, the receiver re-reads until all match.
//   word 0  envelope: value = [31:16] context, [15:0] tag (pack_env)
//   word 1  info:     value = [31:24] SlotFlags, [23:0] message bytes
//   eager:       words 2.. payload, 4 bytes per word (little-endian)
//   rendezvous:  words 2,3 sender buffer (lo, hi); 4,5 done-word address
//                (lo, hi); 6 value for the done word; 7 message bytes >> 24
enum SlotField { SF_ENV = 0, SF_INFO = 1, SF_PAYLOAD = 2, SF_SRC_LO = 2, SF_SRC_HI = 3, SF_DONE_LO = 4,
                 SF_DONE_HI = 5, SF_DONE_VAL = 6, SF_BYTES_HI = 7, SF_RNDV_WORDS = 8 };
constexpr int kMaxEagerBytes = (1 << 24) - 1;
__host__ __device__ inline uint64_t slot_word(uint32_t seq1, uint32_t value) {
    return ((uint64_t)seq1 << 32) | value;
}
__host__ __device__ inline uint32_t slot_word_seq(uint64_t w) { return (uint32_t)(w >> 32); }
__host__ __device__ inline uint32_t slot_word_value(uint64_t w) { return (uint32_t)w; }
// Words a message of `bytes` occupies.
__host__ __device__ inline int eager_words(uint64_t bytes) { return SF_PAYLOAD + (int)((bytes + 3) / 4); }
// Bytes of one slot for an eager limit (whole 64-byte lines).
__host__ __device__ inline int slot_bytes_for(int eager_bytes) {
    int words = eager_words((uint64_t)eager_bytes);
    if (words < SF_RNDV_WORDS) words = SF_RNDV_WORDS;
    return (words * 8 + 63) / 64 * 64;
}

// Per-member entry of a communicator's collective table.
struct alignas(64) CollSlot {
    uint64_t w[kMaxCommSlots];
};

// Communicator state shared by all members (HBM).  MPI_Comm handles other
// than the predefined constants point at one of these.
struct alignas(128) CommShared {
    uint64_t bar;        // monotonically counting barrier (own 128B line)
    uint64_t pad0[15];
    int ctx;             // context id carried by every message
    int size;
    int kind;            // 0 world, 1 self, 2 derived
    int pad1;
    int* world_of;       // [size]: comm rank -> world rank
    int* rank_of_world;  // [world size]: world rank -> comm rank or -1
    CollSlot* slots;     // [size]: collective pointer table
    uint64_t pool_bytes; // derived: size of its block in the communicator pool
};

// The communicator pool (World::comm_pool) starts with its allocator's state:
// a lock, the list of blocks MPI_Comm_free returned, and at kPoolCtxBitsOff a
// bitmap of the context ids it returned (32768 bits).  Blocks follow.
struct PoolHead {
    int lock;
    int pad;
    uint64_t free_head;  // offset of the first free block, 0 for none
};
struct PoolFree {
    uint64_t bytes;
    uint64_t next;  // offset of the next free block, 0 at the end
};
constexpr uint64_t kPoolCtxBitsOff = 4096;
constexpr uint64_t kPoolHeadBytes = 8192;
// Context ids are 15-bit (3 .. kMaxCtx; 0-2 are NULL, WORLD and SELF).  A
// build may lower the limit (GPUMPI_MAX_CTX) to exercise id reuse quickly.
#ifndef GPUMPI_MAX_CTX
#define GPUMPI_MAX_CTX 32767
#endif
constexpr int kMaxCtx = GPUMPI_MAX_CTX;
static_assert(kMaxCtx >= 3 && kMaxCtx <= 32767 && (kMaxCtx + 1) % 32 == 0, "context id limit");

struct ErrorRecord {
    int code;      // first error's MPI error code
    int rank;      // world rank that raised it
    int line;      // source line in the runtime
    int count;     // number of errors raised
    char msg[112];
};

// Device -> host service slot of one rank, in pinned coherent host memory:
// file I/O (MPI_File_*) and formatted text output, carried out by the host
// while the launch runs (the reference's delegate_to_host ring,
// gpu_libs/gpu_libc/cuda_mpi.cuh:392-435 and gpu_main/libc_processor.cu:15-84).
// A rank has at most one request outstanding: its program waits for each.
//   device: fill op/args/data, release, state = POSTED, wait for DONE
//   host:   acquire POSTED, carry out, write results, release, state = DONE
//   device: read results, state = IDLE
constexpr int kSvcSlotBytes = 16384;
constexpr int kSvcDataBytes = kSvcSlotBytes - 64;
enum SvcState : uint32_t { SVC_IDLE = 0, SVC_POSTED = 1, SVC_DONE = 2 };
enum SvcOp : int32_t {
    SVC_OPEN = 1,    // data: path; i[0] amode -> l[0] handle, l[1] size, i[1] error
    SVC_CLOSE = 2,   // l[0] handle, i[0] delete on close -> i[1] error
    SVC_READ = 3,    // l[0] handle, l[1] byte offset, i[0] bytes -> data, i[1] bytes read, i[2] error
    SVC_WRITE = 4,   // l[0] handle, l[1] byte offset, i[0] bytes, data -> i[1] bytes written, i[2] error
    SVC_SIZE = 5,    // l[0] handle -> l[1] size, i[1] error
    SVC_DELETE = 6,  // data: path -> i[1] error
    SVC_SYNC = 7,    // l[0] handle -> i[1] error
    SVC_FORMAT = 8,  // data: format, then argument records; i[0] stream (0: back into data,
                     // 1 stdout, 2 stderr), i[1] argument count, i[2] offset of the records,
                     // i[3] capacity for stream 0 -> i[4] formatted length
};
// One printf argument as the device packs it (strings follow inline, padded
// to 8 bytes).
enum SvcArgKind : int32_t { SVC_ARG_INT = 1, SVC_ARG_UINT = 2, SVC_ARG_DOUBLE = 3, SVC_ARG_STR = 4, SVC_ARG_PTR = 5 };
struct SvcArg {
    int32_t kind;
    int32_t len;  // string bytes that follow (without the NUL)
    uint64_t bits;
};
struct alignas(64) SvcSlot {
    uint32_t state;
    int32_t op;
    int32_t i[6];
    int64_t l[4];
    uint8_t data[kSvcDataBytes];
};
static_assert(sizeof(SvcSlot) == kSvcSlotBytes, "service slot layout");

struct World {
    Config cfg;
    uint8_t* slots;        // [P*P*Q] slots of slot_bytes, index (dst*P+src)*Q+k
    uint32_t* chan_head;   // [P*P] receiver-published consumed count, dst*P+src
    uint32_t* done_words;  // [P*R] rendezvous completion words
    uint8_t* staging;      // [P*staging_bytes]
    uint8_t* arena;        // [P*arena_bytes]
    uint64_t* arena_top;   // [P] bump offsets
    CommShared* comm_world;
    CommShared* comm_self; // [P]
    uint8_t* comm_pool;    // bump region for derived communicators
    uint64_t* comm_pool_top;
    int* next_ctx;
    int* rank_ret;         // [P] return value of each rank's program
    uint64_t* rank_stats;  // [P*8] per-rank counters
    int* abort_word;       // pinned host
    ErrorRecord* err;      // pinned host
    void* user_args;       // program arguments (opaque)
    // Channel cursors and request stamps survive between launches of the same
    // world shape, so a launch needs no reset of the P*P*Q slots: each rank
    // restores them at start and saves them at exit.
    uint8_t* saves;        // [P * save_bytes]
    uint64_t save_bytes;
    // Collective member tables (two words per member) when they do not fit
    // in LDS beside the other per-peer arrays: [P * 2P], rank-private rows.
    uint64_t* members_hbm;
    // Device copy of the abort flag: set (release) by a rank that raises an
    // MPI error, read by every barrier after it completes.
    uint32_t* dev_abort;
    // [P] service slots (pinned host), or null when no host serves this
    // launch (then file I/O fails with MPI_ERR_IO and text output is dropped).
    SvcSlot* svc;
};

// Per-rank save block: stamp, validity magic, then RxChan[P] and TxChan[P]
// from byte 64.  A zeroed block (after a host reset) means "start fresh".
constexpr uint32_t kSaveMagic = 0x53415645u;
__host__ __device__ inline uint64_t save_block_bytes(int nranks) {
    return 64 + (uint64_t)nranks * (16 + 8);
}

}  // namespace rt
}  // namespace gpumpi
