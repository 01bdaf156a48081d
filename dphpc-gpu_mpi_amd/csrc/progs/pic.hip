// 2-D particle-in-cell cycle on a px x py grid of ranks (the C5 workload):
// push -> MPI_Sendrecv migration of leaving particles to the 8 neighbours
// (uneven counts) -> NGP density deposition -> MPI_Allreduce of the kinetic
// energy (exact FP64, and FP32 through the BF16 wire) -> two-cell density
// halo exchange -> field.  Every rule (generator, arithmetic order, the
// stable partition, the order arrivals are appended in, the halo strips) is
// the one oracle/oracle.h states for orc_pic_run, so the results compare bit
// for bit.  The reference's PIC (toolchain_tests/sputnipic/src/mpi_comm.c:
// 21-307) decomposes particles with Scatterv/Send/Recv; this one decomposes
// space, which is where Sendrecv with uneven counts comes in.
//
// Program `pic`: the whole cycle, FP64 and BF16-wire energies.
//
// args: [1..4] px py nx ny, [5] particles per rank, [6] cycles, [7] seed,
// [8..10] vth drift qm (FP32 bits), [11] energy out (double[cycles]),
// [12] BF16-wire energy out (float[cycles]), [13] final counts
// (int64[P]) or 0, [14] particle hashes (uint64[P]) or 0, [15] rho hashes
// (uint64[P]) or 0, [16] phase seconds of rank 0 summed over the cycles
// (double[7]: push+partition, particle exchange, deposit+energy fold,
// energy Allreduce, density halo exchange, field, total) or 0, [17] particle
// capacity per rank.
#include "pic.hpp"

GPUMPI_PROGRAM(pic) { return pic_program<true>(argc, argv); }
