// Registry of the rank programs built into libgpumpi_progs.so, so a host
// (tests, benchmarks, ctypes) can launch one by name through the C-ABI
// (gpumpi_prog_lookup + gpumpi_run).
#pragma once

#include <gpumpi/mpi.h>

namespace gpumpi {
namespace progs {

struct Registrar {
    Registrar(const char* name, const void* kernel);
};

}  // namespace progs
}  // namespace gpumpi

// Defines rank program `name`: `__device__ int prog_<name>(int argc, char** argv)`.
#define GPUMPI_PROGRAM(name)                                                                  \
    __device__ int prog_##name(int argc, char** argv);                                        \
    static ::gpumpi::progs::Registrar gpumpi_reg_##name(#name,                                \
                                                        (const void*)&gpumpi_rank_kernel<prog_##name>); \
    __device__ int prog_##name(int argc, char** argv)

// Every program receives a block of 64-bit words (the launch's user_args).
__device__ inline uint64_t* prog_args() { return gpumpi_user_args<uint64_t>(); }
template <typename T>
__device__ inline T* prog_ptr(int i) { return (T*)(uintptr_t)prog_args()[i]; }
__device__ inline int64_t prog_int(int i) { return (int64_t)prog_args()[i]; }
