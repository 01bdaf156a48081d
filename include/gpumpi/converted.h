/*
 * First header of a program the source converter produced
 * (dphpc-gpu_mpi_amd/converter.py): the C headers the program may include
 * (so that their declarations come before the name mappings below), the
 * device MPI API, and the device printf family and C library subset under
 * their standard names.  The reference's converter prepended global_vars.cuh
 * and its gpumpicc driver the gpu_libc remaps (converter.cpp:233,
 * scripts/gpumpicc.py:12-201).
 */
#ifndef GPUMPI_CONVERTED_H
#define GPUMPI_CONVERTED_H

#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpumpi/mpi.h"

#define GPUMPI_STDIO_NAMES
#include "gpumpi/stdio.h"
#define GPUMPI_LIBC_NAMES
#include "gpumpi/libc.h"

#endif /* GPUMPI_CONVERTED_H */
