// MPI-IO and the printf family on the host service channel, as one rank
// program.  The steps restate the reference's IO tests
// (tests/test_io_file_open.cu: amode legality, open/close/delete;
// test_io_file_access.cu: read, read on a write-only or sequential file,
// get_size, views; test_io_file_write.cu: the write's return value) against
// the MPI standard, then cover what those tests do not: explicit offsets,
// the shared file pointer, transfers larger than one service slot, seeking
// from the end, append and exclusive opens, and formatted text.
//
// args: [0] records (int[P * kIoRec]), [1] directory (NUL-terminated, no
// trailing slash), [2] sprintf output (char[P * 256]), [3] N ints per rank
// for the view round trip.
#include "registry.hpp"

#include "gpumpi/stdio.h"

namespace {

constexpr int kIoRec = 64;

// dst = dir + "/" + name (dst holds 256 bytes).
__device__ inline void io_path(char* dst, const char* dir, const char* name) {
    int n = 0;
    for (int i = 0; dir[i] && n < 200; i++) dst[n++] = dir[i];
    dst[n++] = '/';
    for (int i = 0; name[i] && n < 255; i++) dst[n++] = name[i];
    dst[n] = 0;
}

}  // namespace

GPUMPI_PROGRAM(io_scenarios) {
    MPI_Init(&argc, &argv);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    int* rec = prog_ptr<int>(0) + (uint64_t)rank * kIoRec;
    const char* dir = prog_ptr<const char>(1);
    char* txt = prog_ptr<char>(2) + (uint64_t)rank * 256;
    const int N = (int)prog_int(3);
    for (int k = 0; k < kIoRec; k++) rec[k] = -1;
    MPI_Info info = MPI_INFO_NULL;
    MPI_File fh;
    char f1[256], f2[256], f3[256], f4[256], f5[256], fx[256];
    io_path(f1, dir, "io_test_1.txt");
    io_path(f2, dir, "io_test_2.txt");
    io_path(f3, dir, "io_test_view.out");
    io_path(f4, dir, "io_test_shared.out");
    io_path(f5, dir, "io_test_large.out");
    io_path(fx, dir, "io_test_missing.txt");

    // 0: the seven illegal amodes of test_io_file_open.cu:34-96
    const int bad[7] = {MPI_MODE_RDONLY | MPI_MODE_RDWR, MPI_MODE_RDONLY | MPI_MODE_WRONLY,
                        MPI_MODE_RDWR | MPI_MODE_WRONLY, MPI_MODE_RDONLY | MPI_MODE_RDWR | MPI_MODE_WRONLY,
                        MPI_MODE_CREATE | MPI_MODE_RDONLY, MPI_MODE_EXCL | MPI_MODE_RDONLY,
                        MPI_MODE_SEQUENTIAL | MPI_MODE_RDWR};
    int nbad = 0;
    for (int k = 0; k < 7; k++) nbad += MPI_File_open(MPI_COMM_WORLD, f1, bad[k], info, &fh) == MPI_ERR_AMODE;
    rec[0] = nbad;
    // 1-4: open, write from rank 0 (test_io_file_write.cu), close
    const char* hello = "ccchello world";
    rec[1] = MPI_File_open(MPI_COMM_WORLD, f1, MPI_MODE_RDWR | MPI_MODE_CREATE, info, &fh);
    rec[2] = fh.file != nullptr;
    if (rank == 0) rec[3] = MPI_File_write(fh, hello, 14, MPI_CHAR, nullptr);
    rec[4] = MPI_File_close(&fh);
    // 5-9: read back on every rank (test_io_file_access.cu:8-35), size
    rec[5] = MPI_File_open(MPI_COMM_WORLD, f1, MPI_MODE_RDONLY, info, &fh);
    char ch = 0;
    MPI_Status st;
    rec[6] = MPI_File_read(fh, &ch, 1, MPI_CHAR, &st);
    rec[7] = ch;
    MPI_Offset size = -1;
    MPI_File_get_size(fh, &size);
    rec[8] = (int)size;
    int cnt = -1;
    MPI_Get_count(&st, MPI_CHAR, &cnt);
    rec[9] = cnt;
    MPI_File_close(&fh);
    // 10-14: read on a write-only / sequential file (test_io_file_access.cu:42-65)
    rec[10] = MPI_File_open(MPI_COMM_WORLD, f2, MPI_MODE_WRONLY | MPI_MODE_CREATE, info, &fh);
    rec[11] = MPI_File_read(fh, &ch, 1, MPI_CHAR, nullptr);
    MPI_File_close(&fh);
    rec[12] = MPI_File_open(MPI_COMM_WORLD, f2, MPI_MODE_RDONLY | MPI_MODE_SEQUENTIAL, info, &fh);
    rec[13] = MPI_File_read(fh, &ch, 1, MPI_CHAR, nullptr);
    MPI_File_close(&fh);
    MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 0) rec[14] = MPI_File_delete(f2, info);
    MPI_Barrier(MPI_COMM_WORLD);
    // 15-20: views (test_io_file_access.cu:100-157)
    MPI_File_open(MPI_COMM_WORLD, f3, MPI_MODE_RDWR | MPI_MODE_CREATE, info, &fh);
    MPI_Offset disp = -1;
    MPI_Datatype et = MPI_INT, ft = MPI_INT;
    char rep[16] = {0};
    MPI_File_get_view(fh, &disp, &et, &ft, rep);
    rec[15] = (int)disp;
    rec[16] = et == MPI_CHAR && ft == MPI_CHAR;
    rec[17] = rep[0] == 'n' && rep[1] == 'a' && rep[5] == 'e' && rep[6] == 0;
    rec[18] = MPI_File_set_view(fh, (MPI_Offset)rank * N * 4, MPI_INT, MPI_INT, "native", info);
    MPI_File_get_view(fh, &disp, &et, &ft, rep);
    rec[19] = disp == (MPI_Offset)rank * N * 4 && et == MPI_INT && ft == MPI_INT;
    // 20-25: N ints through the view, back again (test_io_file_access.cu:161-215)
    int* wbuf = (int*)gpumpi_malloc(sizeof(int) * (uint64_t)N);
    int* rbuf = (int*)gpumpi_malloc(sizeof(int) * (uint64_t)N);
    for (int i = gpumpi::rt::lane_id(); i < N; i += 64) {
        wbuf[i] = rank * 1000 + i;
        rbuf[i] = -7;
    }
    gpumpi::rt::wave_sync();
    rec[20] = MPI_File_write(fh, wbuf, N, MPI_INT, nullptr);
    MPI_Offset pos = -1;
    MPI_File_get_position(fh, &pos);
    rec[21] = (int)pos;
    rec[22] = MPI_File_seek(fh, 0, MPI_SEEK_SET);
    rec[23] = MPI_File_read(fh, rbuf, N, MPI_INT, nullptr);
    int same = 1;
    for (int i = 0; i < N; i++) same &= rbuf[i] == wbuf[i];
    rec[24] = same;
    rec[25] = MPI_File_seek(fh, -1, MPI_SEEK_SET);  // MPI_ERR_UNSUPPORTED_OPERATION
    MPI_File_close(&fh);
    // 26-29: the next rank's ints through an explicit byte offset, the size
    MPI_File_open(MPI_COMM_WORLD, f3, MPI_MODE_RDONLY, info, &fh);
    const int nx = (rank + 1) % P;
    rec[26] = MPI_File_read_at(fh, (MPI_Offset)nx * N * 4, rbuf, N * 4, MPI_CHAR, nullptr);
    same = 1;
    for (int i = 0; i < N; i++) same &= rbuf[i] == nx * 1000 + i;
    rec[27] = same;
    MPI_File_get_size(fh, &size);
    rec[28] = size == (MPI_Offset)P * N * 4;
    rec[29] = MPI_File_seek(fh, 0, MPI_SEEK_END) == MPI_SUCCESS;
    MPI_File_get_position(fh, &pos);
    rec[30] = pos == (MPI_Offset)P * N * 4;
    MPI_File_close(&fh);
    // 31-34: the shared file pointer: every rank appends a 16-byte record
    MPI_File_open(MPI_COMM_WORLD, f4, MPI_MODE_RDWR | MPI_MODE_CREATE, info, &fh);
    char recb[16];
    for (int k = 0; k < 16; k++) recb[k] = (char)('A' + rank % 26);
    rec[31] = MPI_File_write_shared(fh, recb, 16, MPI_CHAR, nullptr);
    MPI_Barrier(MPI_COMM_WORLD);
    MPI_File_get_position_shared(fh, &pos);
    rec[32] = pos == (MPI_Offset)P * 16;
    MPI_File_close(&fh);
    if (rank == 0) {
        MPI_File f;
        MPI_File_open(MPI_COMM_SELF, f4, MPI_MODE_RDONLY, info, &f);
        char* all = (char*)gpumpi_malloc((uint64_t)P * 16);
        const int got = MPI_File_read(f, all, P * 16, MPI_CHAR, nullptr);
        int ok = got == P * 16;
        // every 16-byte record is whole and the letters appear as often as
        // ranks map to them
        for (int r = 0; r < P && ok; r++)
            for (int k = 1; k < 16; k++) ok &= all[r * 16 + k] == all[r * 16];
        for (int l = 0; l < 26 && ok; l++) {
            int want = 0, have = 0;
            for (int r = 0; r < P; r++) want += r % 26 == l;
            for (int r = 0; r < P; r++) have += all[r * 16] == 'A' + l;
            ok &= want == have;
        }
        rec[33] = ok;
        MPI_File_close(&f);
        gpumpi_free(all, (uint64_t)P * 16);
    }
    MPI_Barrier(MPI_COMM_WORLD);
    // 34-35: append and exclusive opens, a missing file
    MPI_File_open(MPI_COMM_WORLD, f4, MPI_MODE_WRONLY | MPI_MODE_APPEND, info, &fh);
    MPI_File_get_position(fh, &pos);
    rec[34] = pos == (MPI_Offset)P * 16;
    MPI_File_close(&fh);
    rec[35] = MPI_File_open(MPI_COMM_WORLD, f4, MPI_MODE_RDWR | MPI_MODE_CREATE | MPI_MODE_EXCL, info, &fh);
    rec[36] = MPI_File_open(MPI_COMM_WORLD, fx, MPI_MODE_RDONLY, info, &fh);
    // 37-38: 40000 bytes (three service slots) from rank 0, back again
    if (rank == 0) {
        MPI_File f;
        MPI_File_open(MPI_COMM_SELF, f5, MPI_MODE_RDWR | MPI_MODE_CREATE | MPI_MODE_DELETE_ON_CLOSE, info, &f);
        const int nb = 40000;
        uint8_t* big = (uint8_t*)gpumpi_malloc(nb);
        uint8_t* back = (uint8_t*)gpumpi_malloc(nb);
        for (int i = gpumpi::rt::lane_id(); i < nb; i += 64) {
            big[i] = (uint8_t)(i * 7 + 3);
            back[i] = 0;
        }
        gpumpi::rt::wave_sync();
        rec[37] = MPI_File_write_at(f, 0, big, nb, MPI_BYTE, nullptr);
        const int got = MPI_File_read_at(f, 0, back, nb, MPI_BYTE, nullptr);
        int ok = got == nb;
        for (int i = 0; i < nb && ok; i++) ok &= back[i] == big[i];
        rec[38] = ok;
        MPI_File_close(&f);  // deletes it
        gpumpi_free(back, nb);
        gpumpi_free(big, nb);
    }
    // 39-41: formatted text
    rec[39] = gpumpi_sprintf(txt, "%d|%5.2f|%s|%x|%lld|%c|%%|%-4d|%e|%08.3f|%+i|%u", rank, 3.14159, "abc", 255,
                             (long long)1 << 40, 'Z', 7, 12345.678, -2.5, rank, 4000000000u);
    char small[8];
    rec[40] = gpumpi_snprintf(small, 5, "%d", 1234567);
    rec[41] = small[0] == '1' && small[3] == '4' && small[4] == 0;
    rec[42] = gpumpi_printf("io_scenarios: rank %d of %d\n", rank, P);
    MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 0) {
        rec[43] = MPI_File_delete(f1, info);
        rec[44] = MPI_File_delete(f3, info);
        rec[45] = MPI_File_delete(f4, info);
        rec[46] = MPI_File_delete(fx, info);  // MPI_ERR_NO_SUCH_FILE
    }
    gpumpi_free(rbuf, sizeof(int) * (uint64_t)N);
    gpumpi_free(wbuf, sizeof(int) * (uint64_t)N);
    MPI_Finalize();
    return 0;
}
