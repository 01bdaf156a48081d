// The reference's toolchain_tests/pi/cpi.c after conversion: main renamed to
// __gpu_main, built with include/gpumpi/gpu_main.h into a standalone
// executable that takes the reference's `---gpumpi` options.
#include <math.h>

#include "mpi.h"

__device__ double f(double a) { return (4.0 / (1.0 + a * a)); }

__device__ int __gpu_main(int argc, char* argv[]) {
    int n, myid, numprocs, i;
    double PI25DT = 3.141592653589793238462643;
    double mypi, pi = 0.0, h, sum, x;
    double startwtime = 0.0, endwtime;
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &numprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    n = 10000;
    if (argc >= 2) {
        n = 0;
        for (const char* p = argv[1]; *p >= '0' && *p <= '9'; p++) n = n * 10 + (*p - '0');
    }
    if (myid == 0) startwtime = MPI_Wtime();
    MPI_Bcast(&n, 1, MPI_INT, 0, MPI_COMM_WORLD);
    h = 1.0 / (double)n;
    sum = 0.0;
    for (i = myid + 1; i <= n; i += numprocs) {
        x = h * ((double)i - 0.5);
        sum += f(x);
    }
    mypi = h * sum;
    MPI_Reduce(&mypi, &pi, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    if (myid == 0) {
        endwtime = MPI_Wtime();
        if (gpumpi_is_leader_lane())
            printf("pi is approximately %.16f, Error is %.16f\nwall clock time = %f\n", pi, fabs(pi - PI25DT),
                   endwtime - startwtime);
    }
    MPI_Finalize();
    return 0;
}

#include "gpumpi/gpu_main.h"
