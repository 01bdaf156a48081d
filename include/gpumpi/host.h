/*
 * gpumpi host C-ABI: launch a device MPI program (one workgroup per rank)
 * on one MI355X.  Plain C types only, so it can be bound from C, C++,
 * Python ctypes or any other FFI.
 *
 * Replaces the reference's host launcher libgpu_main.a
 * (gpu_libs/gpu_main/main.cu:133-321: `---gpumpi` CLI, occupancy checks,
 * heap/stack limits, kernel launch, host polling loop, return-code
 * aggregation) and the test launcher tests/test_runner.cuh:26-86.
 */
#ifndef GPUMPI_HOST_H
#define GPUMPI_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Launch configuration.  main.cu:70-78 options map as:
 *   -n numProcs          -> num_procs
 *   -b threadsPerBlock   -> threads_per_rank (a rank is a workgroup here)
 *   -g blocksPerGrid     -> accepted; must equal num_procs if given
 *   -s stackSize         -> stack_size (bytes per lane)
 *   -p heapSize          -> heap_bytes (per-rank device heap * num_procs)
 *   -e pendingBufferSize -> max_requests (per-rank request pool)
 * New knobs: chan_slots, eager_bytes, helper_min_bytes, staging_bytes,
 * xcd_map, timeout_s. */
typedef struct gpumpi_config {
    int num_procs;
    int threads_per_rank;
    int blocks_per_grid;
    int max_requests;
    int chan_slots;
    int eager_bytes;
    int helper_min_bytes;
    int xcd_map;
    size_t stack_size;
    size_t heap_bytes;
    size_t staging_bytes;
    size_t comm_pool_bytes;
    double timeout_s;
    int verbose;
    int user_lds_bytes;   /* LDS per rank for the program (gpumpi_user_lds) */
    int reserved[6];
} gpumpi_config;

/* What a launch reports back. */
typedef struct gpumpi_result {
    int status;           /* 0 ok, otherwise a GPUMPI_E* code */
    int any_rank_failed;  /* 1 if a rank's program returned non-zero (main.cu:127-129) */
    int mpi_error;        /* first MPI error raised on the device, 0 if none */
    int mpi_error_rank;   /* rank that raised it */
    int mpi_error_line;   /* runtime source line */
    int mpi_error_count;  /* errors raised in total */
    int timed_out;        /* watchdog fired */
    int lds_bytes;        /* dynamic LDS per rank workgroup */
    double kernel_ms;     /* launch duration, HIP events on the launch stream */
    char mpi_error_msg[112];
} gpumpi_result;

enum {
    GPUMPI_OK = 0,
    GPUMPI_EINVAL = 1,     /* bad configuration */
    GPUMPI_EHIP = 2,       /* a HIP runtime call failed */
    GPUMPI_ETOOBIG = 3,    /* ranks cannot all be resident (grid barriers would deadlock) */
    GPUMPI_EDEVICE = 4,    /* the device program raised an MPI error */
    GPUMPI_ETIMEOUT = 5,   /* watchdog aborted the launch */
    GPUMPI_ENOMEM = 6
};

/* Fill `cfg` with defaults for `num_procs` ranks. */
int gpumpi_default_config(gpumpi_config* cfg, int num_procs);

/* Parse the reference's CLI: everything after "---gpumpi" (main.cu:48-97).
 * Returns the index of "---gpumpi" (the program's own argc), or -1 if the
 * marker is missing, -2 on a parse error; `cfg` is updated in place. */
int gpumpi_parse_args(int argc, char** argv, gpumpi_config* cfg);

/* Apply a tuning file written by the autotuner (dphpc-gpu_mpi_amd/gpumpi/
 * tune.py): "key value" lines for threads_per_rank, helper_min_bytes,
 * eager_bytes and chan_slots.  `cfg` changes only if every line is valid;
 * returns GPUMPI_OK or GPUMPI_EINVAL (gpumpi_last_error says why).  Also
 * reachable as `---gpumpi --tuning FILE`.  No counterpart in the reference:
 * its launch knobs are the fixed CLI options of main.cu:70-78. */
int gpumpi_apply_tuning(gpumpi_config* cfg, const char* path);

/* Launch `kernel` (a gpumpi_rank_kernel<main> instance) with cfg->num_procs
 * ranks and wait for it.  `argc/argv` are copied to device memory and given
 * to every rank's program; `user_args` (a device pointer or NULL) is
 * returned by gpumpi_user_args<T>() on the device.  Blocks until the kernel
 * ends or the watchdog fires; returns result->status. */
int gpumpi_run(const gpumpi_config* cfg, const void* kernel, int argc, char** argv, void* user_args,
               gpumpi_result* result);

/* Capture one launch of `kernel` into a hipGraph and replay it `replays`
 * times (iteration loops of a whole MPI program without host launch work).
 * Channel state carries over between replays.  replay_ms[i] (optional)
 * receives each replay's HIP-event time; result->kernel_ms their mean. */
int gpumpi_run_graph(const gpumpi_config* cfg, const void* kernel, int argc, char** argv, void* user_args,
                     int replays, gpumpi_result* result, double* replay_ms);

/* gpumpi_run without any allocation -- no counterpart in the reference; for
 * callers that cannot allocate (tight serving loops, code that must not call
 * hipMalloc).  The whole world (channels, staging, rank heaps, communicator
 * tables, error record, argv) lives in one static device array of 64 MiB per
 * GPU, so configurations whose world does not fit are refused
 * (GPUMPI_EINVAL, with the bytes needed in gpumpi_last_error()); use small
 * staging_bytes / heap_bytes / comm_pool_bytes.  Every call starts from fresh
 * channel state; it runs on the null stream with no watchdog (timeout_s is
 * ignored; kernel_ms is host wall time).  Calls that overlap on one GPU
 * share the array and race. */
int gpumpi_run_noalloc(const gpumpi_config* cfg, const void* kernel, int argc, char** argv, void* user_args,
                       gpumpi_result* result);

/* Drop-in host main: parse `---gpumpi`, run, return 0 or 1 like main.cu. */
int gpumpi_main(int argc, char** argv, const void* kernel);

/* Device memory for hosts that hold no HIP context of their own (tests,
 * FFI callers).  All return 0 on success. */
int gpumpi_device_alloc(void** ptr, size_t bytes);
int gpumpi_device_free(void* ptr);
int gpumpi_memcpy_htod(void* dst, const void* src, size_t bytes);
int gpumpi_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int gpumpi_memset(void* dst, int value, size_t bytes);

/* Carry out the pending requests of the running launch's ranks on the host
 * service channel (MPI_File_*, the printf family) once; returns how many.
 * gpumpi_run and gpumpi_run_graph do this themselves while they wait; a host
 * that waits for a launch by other means calls it in its loop.  Replaces the
 * reference's host polling of its device-to-host queue
 * (gpu_libs/gpu_libc/cuda_mpi.cuh:414-427, gpu_main/main.cu:260-300). */
int gpumpi_service_poll(void);

/* Start the next launch from fresh channel state (by default channel
 * cursors carry over between clean launches of the same world shape). */
int gpumpi_reset_world(void);

/* Free the device world kept between launches (channels, heaps, staging,
 * stream); the next launch allocates it again.  For hosts that finish with
 * the runtime before the process ends.  No counterpart in the reference,
 * whose host main() exits after its one launch (main.cu:133-321). */
int gpumpi_release_world(void);

/* Largest rank count that can be co-resident for a configuration. */
int gpumpi_max_ranks(const gpumpi_config* cfg, const void* kernel);

/* Text of the last host-side error. */
const char* gpumpi_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GPUMPI_HOST_H */
