"""Expected records of the io_scenarios rank program (csrc/progs/io.hip),
shared by its CPU-simulation test and its GPU test.  Codes are the MPI-IO
error classes of include/gpumpi/mpi.h (the reference's io.cuh:28-44); the
sprintf text is what the host C library's snprintf gives for the same
format and arguments."""
import ctypes

MPI_ERR_AMODE = 3
MPI_ERR_UNSUPPORTED_OPERATION = 5
MPI_ERR_NO_SUCH_FILE = 6
MPI_ERR_FILE_EXISTS = 7
KREC = 64
FMT = b"%d|%5.2f|%s|%x|%lld|%c|%%|%-4d|%e|%08.3f|%+i|%u"


def c_sprintf(rank):
    libc = ctypes.CDLL(None)
    buf = ctypes.create_string_buffer(256)
    n = libc.snprintf(buf, 256, FMT, ctypes.c_int(rank), ctypes.c_double(3.14159), b"abc", ctypes.c_int(255),
                      ctypes.c_longlong(1 << 40), ctypes.c_int(ord("Z")), ctypes.c_int(7), ctypes.c_double(12345.678),
                      ctypes.c_double(-2.5), ctypes.c_int(rank), ctypes.c_uint(4000000000))
    return n, buf.value.decode()


def expected(rank, P, N):
    """{record index: value} for one rank."""
    e = {0: 7, 1: 0, 2: 1, 3: 14 if rank == 0 else -1, 4: 0,
         5: 0, 6: 1, 7: ord("c"), 8: 14, 9: 1,
         10: 0, 11: MPI_ERR_AMODE, 12: 0, 13: MPI_ERR_UNSUPPORTED_OPERATION, 14: 0 if rank == 0 else -1,
         15: 0, 16: 1, 17: 1, 18: 0, 19: 1,
         20: N, 21: N, 22: 0, 23: N, 24: 1, 25: MPI_ERR_UNSUPPORTED_OPERATION,
         26: 4 * N, 27: 1, 28: 1, 29: 1, 30: 1,
         31: 16, 32: 1, 33: 1 if rank == 0 else -1, 34: 1, 35: MPI_ERR_FILE_EXISTS, 36: MPI_ERR_NO_SUCH_FILE,
         37: 40000 if rank == 0 else -1, 38: 1 if rank == 0 else -1,
         39: c_sprintf(rank)[0], 40: 7, 41: 1, 42: len(f"io_scenarios: rank {rank} of {P}\n")}
    for k in (43, 44, 45):
        e[k] = 0 if rank == 0 else -1
    e[46] = MPI_ERR_NO_SUCH_FILE if rank == 0 else -1
    return e


def check(records, texts, P, N):
    """records: int array [P, KREC]; texts: list of P strings."""
    for r in range(P):
        for k, v in expected(r, P, N).items():
            assert records[r][k] == v, f"rank {r} record {k}: {records[r][k]} != {v}"
        assert texts[r] == c_sprintf(r)[1], (r, texts[r])
