/*
 * `#include "mpi.h"` of a converted MPI program resolves here and gets the
 * device MPI API (the reference remapped <mpi.h> to gpu_libs/gpu_mpi/
 * mpi.h.cuh, scripts/gpumpicc.py:177-182).
 */
#ifndef GPUMPI_MPI_SHIM_H
#define GPUMPI_MPI_SHIM_H
#include "gpumpi/mpi.h"
#endif
