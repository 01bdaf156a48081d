/*
 * Host entry of a converted program.  Include this header once, in the
 * translation unit that defines `__device__ int __gpu_main(int, char**)`;
 * it instantiates the rank kernel for that program and defines host main(),
 * which keeps the reference's `---gpumpi` command line
 * (gpu_libs/gpu_main/main.cu:133-321, which linked __gpu_main through
 * separable compilation instead).
 *
 *   hipcc --offload-arch=gfx950 -ffp-contract=off -I<repo>/include \
 *         -I<repo>/dphpc-gpu_mpi_amd/csrc prog.cpp -L<lib> -lgpumpi -o prog
 *   ./prog ARGS ---gpumpi -n 64 -b 256
 */
#ifndef GPUMPI_GPU_MAIN_H
#define GPUMPI_GPU_MAIN_H

#include "gpumpi/host.h"
#include "gpumpi/mpi.h"

int main(int argc, char** argv) {
    return gpumpi_main(argc, argv, (const void*)&gpumpi_rank_kernel<__gpu_main>);
}

#endif
