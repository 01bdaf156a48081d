// Host build of the reduction fold (ops.hpp fold_builtin), callable through
// ctypes: runs every lane of a wave over the slice [lo, hi) exactly as the
// device's CMD_REDUCE does, so the runtime's own arithmetic can be compared
// with the oracle on the CPU.
#include "gpumpi/mpi.h"

#include <vector>

thread_local int gpumpi_host_lane = 0;
thread_local gpumpi_host_idx threadIdx, blockIdx, blockDim, gridDim;
namespace gpumpi {
namespace rt {
thread_local uint8_t (*gpumpi_host_lds)[] = nullptr;
}
}  // namespace gpumpi

using namespace gpumpi::rt;

extern "C" int host_fold(int dt, int op, int nmembers, const void* const* ins, void* const* outs, uint64_t lo,
                         uint64_t hi, int root, int vec, int nt) {
    if (dt < 0 || op < 0 || op >= GPUMPI_NUM_BUILTIN_OPS || !op_allowed((MPI_Datatype)dt, op)) return -1;
    std::vector<CollSlot> tab((size_t)nmembers);
    for (int r = 0; r < nmembers; r++) {
        std::memset(&tab[r], 0, sizeof(CollSlot));
        tab[r].w[SW_SBUF] = (uint64_t)(uintptr_t)ins[r];
        tab[r].w[SW_RBUF] = (uint64_t)(uintptr_t)outs[r];
    }
    FoldArgs A;
    A.ins_w = &tab[0].w[SW_SBUF];
    A.outs_w = &tab[0].w[SW_RBUF];
    A.stride_words = kMaxCommSlots;
    A.nmembers = nmembers;
    A.root = root;
    A.vec = vec;
    A.lo = lo;
    A.hi = hi;
    for (int t = 0; t < nt; t++) {
        gpumpi_host_lane = t & 63;
        fold_builtin(dt, op, A, t, nt);
    }
    return 0;
}
