"""The source converter (dphpc-gpu_mpi_amd/converter.py) on the CPU: the C
view of the device MPI header it parses programs against, and the rewrites
of a C MPI program into a rank program (reference:
source_converter/converter.cpp:27-313).  Running converted programs is in
tests/test_simt.py (CPU simulation) and tests/test_gpu_y_main.py (GPU)."""
import os
import re
import subprocess

import pytest

import converter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RING = os.path.join(ROOT, "dphpc-gpu_mpi_amd", "examples", "ring.c")


def test_c_view_declares_every_device_mpi_function(tmp_path):
    text = open(os.path.join(ROOT, "include", "gpumpi", "mpi.h")).read()
    names = set(re.findall(r"^__device__ (?:int|double) (MPI_\w+)\(", text, re.M))
    assert len(names) >= 80
    c_h = converter.c_view_of_mpi_h(text)
    for n in names:
        assert re.search(rf"^(?:int|double) {n}\(", c_h, re.M), n
    (tmp_path / "mpi.h").write_text(c_h)
    src = tmp_path / "use.c"
    src.write_text('#include "mpi.h"\nint main(void) { MPI_Status s; MPI_File f = MPI_FILE_NULL; '
                   'MPI_Info i = MPI_INFO_NULL; MPI_Datatype d = MPI_DOUBLE; (void)s; (void)f; (void)i; (void)d;\n'
                   'return MPI_Init(0, 0); }\n')
    p = subprocess.run([converter.CLANG, "-fsyntax-only", "-x", "c", "-std=gnu11", "-I", str(tmp_path), str(src)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_ring_rewrites(tmp_path):
    out = tmp_path / "ring.cpp"
    assert converter.convert(RING, str(out)) is True
    t = out.read_text()
    assert t.startswith("// Converted from ring.c") and '#include "gpumpi/converted.h"' in t
    assert t.rstrip().endswith('#include "gpumpi/gpu_main.h"')
    assert "__device__ int __gpu_main(int argc, char** argv)" in t
    assert "__device__ static int next_of(int r, int n);" in t  # the prototype too
    assert f"__device__ int hops[{converter.MAX_RANKS}];" in t
    assert f"static int calls[{converter.MAX_RANKS}] = {{0, 0," in t  # static local, per rank
    assert "__gpu_global(hops)++" in t and "++__gpu_global(calls)" in t
    assert "__device__ const int SCALE = 10;" in t and "__gpu_global(SCALE)" not in t  # read-only: one copy
    assert t.count("{7, 11}") == converter.MAX_RANKS  # initialiser per rank
    assert "int __decl_class = 0;" in t and "__decl_class += count_calls();" in t
    assert "(int *)malloc(" in t and "(int *)calloc(" in t  # void* -> int*
    assert re.search(r"MPI_Allreduce\(\(const void \*\)&__gpu_global\(checksum\)\[0\]", t)


def test_a_file_without_main_gets_no_host_entry(tmp_path):
    src = tmp_path / "lib.c"
    src.write_text("int counter;\nint bump(int by) { counter += by; return counter; }\n")
    out = tmp_path / "lib.cpp"
    assert converter.convert(str(src), str(out)) is False
    t = out.read_text()
    assert "gpumpi/gpu_main.h" not in t
    assert "__device__ int bump(int by) { __gpu_global(counter) += by; return __gpu_global(counter); }" in t


def test_parse_errors_are_reported(tmp_path):
    src = tmp_path / "bad.c"
    src.write_text("int main(void) { return undefined_thing; }\n")
    with pytest.raises(RuntimeError, match="parsing"):
        converter.convert(str(src), str(tmp_path / "bad.cpp"))


def test_old_style_definitions_become_prototypes(tmp_path):
    # K&R definitions (NPB's common/randdp.c is written so) get a prototype
    # parameter list; an array parameter is its adjusted pointer type, as C
    # defines it.
    src = tmp_path / "kr.c"
    src.write_text("double scale(x, n, v)\ndouble *x;\nint n;\ndouble v[4];\n{ return x[n] * v[0]; }\n"
                   "int plain(int a) { return a; }\n")
    out = tmp_path / "kr.cpp"
    converter.convert(str(src), str(out))
    t = out.read_text()
    assert "__device__ double scale(double * x, int n, double * v) {" in t
    assert "__device__ int plain(int a) { return a; }" in t


def test_prototypes_of_own_headers_get_device_redeclarations(tmp_path):
    # A program's own header is not rewritten; the converted file redeclares
    # its functions for the device after the last #include, so calls into
    # another converted file link (NPB's c_timers.h).
    inc = tmp_path / "inc"
    inc.mkdir()
    (inc / "timers.h").write_text("#ifndef T_H\n#define T_H\nextern void   timer_start( int n );\n"
                                  "double timer_read(int n);\n#endif\n")
    src = tmp_path / "use.c"
    src.write_text('#include <stdio.h>\n#include "timers.h"\nint f(void) { timer_start(1); '
                   'return (int)timer_read(1); }\n')
    out = tmp_path / "use.cpp"
    converter.convert(str(src), str(out), [str(inc)])
    t = out.read_text()
    body = t.index("__device__ int f(void)")
    for d in ("__device__ extern void timer_start( int n );", "__device__ double timer_read(int n);"):
        assert d in t and t.index(d) < body, d
    assert "printf" not in t.split('#include "timers.h"')[1].split("__device__ int f")[0]  # no libc redeclarations
