"""Converted C programs as standalone executables on the GPU: host main()
of include/gpumpi/gpu_main.h (gpumpi_main) with the reference's `---gpumpi`
command line (gpu_libs/gpu_main/main.cu:48-97), text through the host
service channel.  lib/ring is dphpc-gpu_mpi_amd/examples/ring.c, lib/cpi the
reference's toolchain_tests/pi/cpi.c, lib/is.S.x its NPB IS,
lib/dt.S.x its NPB DT, all built by converter.py and gpumpicc.py
(dphpc-gpu_mpi_amd/build.py)."""
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "dphpc-gpu_mpi_amd", "lib")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _run(args):
    print("[gpu] " + " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)), flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("P", [4, 64])
def test_converted_ring_matches_host_mpi(P):
    exe = os.path.join(LIB, "ring")
    if not os.path.exists(exe):
        pytest.skip("lib/ring is not built")
    p = _run([exe, "5", "---gpumpi", "-n", str(P), "-b", "64"])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    gold = json.load(open(os.path.join(GOLDEN, "ring_mpich.json")))
    assert gold[str(P)] in p.stdout, (gold[str(P)], p.stdout[-2000:])


def test_converted_reference_cpi(orc):
    exe = os.path.join(LIB, "cpi")
    if not os.path.exists(exe):
        pytest.skip("lib/cpi is not built (the reference's cpi.c was not at hand)")
    P, n = 64, 100000000
    p = _run([exe, str(n), "---gpumpi", "-n", str(P), "-b", "256"])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    for r in range(P):
        assert f"Process {r} of {P} is on" in p.stdout
    pi = float(re.search(r"pi is approximately ([0-9.]+)", p.stdout).group(1))
    want, _ = orc.pi_fold(n, P)
    assert pi == float(f"{want:.16f}")
    host = float(json.load(open(os.path.join(GOLDEN, "cpi_mpich.json")))["pi_printed"]["64"])
    assert abs(pi - host) <= 1e-13 * abs(host)


@pytest.mark.parametrize("cls,P,heap", [("S", 1, 64 << 20), ("S", 4, 256 << 20), ("S", 16, 1 << 30)])
def test_converted_reference_npb_is(orc, cls, P, heap):
    # The reference's NPB IS (is.c, c_print_results.c, c_timers.c, unmodified)
    # converted into lib/is.<class>.x: NPB's own verification passes, as it
    # does for is.c under host MPI and for the oracle's restatement.
    exe = os.path.join(LIB, f"is.{cls}.x")
    if not os.path.exists(exe):
        pytest.skip(f"lib/is.{cls}.x is not built (the reference's NPB IS was not at hand)")
    p = _run([exe, "---gpumpi", "-n", str(P), "-b", "256", "-p", str(heap)])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert re.search(r"Verification\s+=\s+SUCCESSFUL", p.stdout), p.stdout[-2000:]
    assert re.search(rf"Total processes\s+=\s+{P}\b", p.stdout)
    res, _, _ = orc.is_run(cls, P)
    assert res.passed_verification == 5 * 10 + P


@pytest.mark.parametrize("graph", ["BH", "WH", "SH"])
def test_converted_reference_npb_dt(graph):
    # The reference's NPB DT class S (dt.c, DGraph.c and the common files,
    # unmodified) converted into lib/dt.S.x: the L2 norm host MPICH prints for
    # the same graph and rank count (tests/golden/npb_dt_mpich.json), and
    # NPB's verification.
    exe = os.path.join(LIB, "dt.S.x")
    if not os.path.exists(exe):
        pytest.skip("lib/dt.S.x is not built (the reference's NPB DT was not at hand)")
    gold = json.load(open(os.path.join(GOLDEN, "npb_dt_mpich.json")))["runs"][graph]
    p = _run([exe, graph, "---gpumpi", "-n", str(gold["ranks"]), "-b", "256", "-p", str(256 << 20)])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    text = p.stdout + p.stderr
    assert f"DT_{graph}.S L2 Norm = {gold['l2_norm']}" in text, text[-2000:]
    assert re.search(r"Verification\s+=\s+SUCCESSFUL", text)


def test_converted_reference_global_var_test():
    # toolchain_tests/global_var (the reference's converter test, 4 ranks as in
    # its test.sh): per-rank globals; non-zero return and "error" lines on any
    # mismatch.
    exe = os.path.join(LIB, "global_var")
    if not os.path.exists(exe):
        pytest.skip("lib/global_var is not built (the reference's toolchain tests were not at hand)")
    p = _run([exe, "---gpumpi", "-n", "4", "-b", "64"])
    assert p.returncode == 0 and "error" not in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]


def test_converted_reference_basic_io_test(tmp_path):
    # toolchain_tests/basicIO/basicIO.c: rank 0's greeting in test.out.
    exe = os.path.join(LIB, "basicIO")
    if not os.path.exists(exe):
        pytest.skip("lib/basicIO is not built (the reference's toolchain tests were not at hand)")
    print("[gpu] basicIO", flush=True)
    p = subprocess.run([exe, "---gpumpi", "-n", "4", "-b", "64"], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert (tmp_path / "test.out").read_bytes() == b"hello world from GPUMPI IO"


def test_converted_reference_contiguous_test():
    # toolchain_tests/datatype/contiguous.c: MPI_Type_contiguous(100, MPI_CHAR).
    exe = os.path.join(LIB, "contiguous")
    if not os.path.exists(exe):
        pytest.skip("lib/contiguous is not built (the reference's toolchain tests were not at hand)")
    p = _run([exe, "---gpumpi", "-n", "2", "-b", "64"])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]


def test_converted_reference_file_view_test(tmp_path):
    # toolchain_tests/basicIO/file_view1.c: 100 ints per rank at rank*400 B.
    import numpy as np
    exe = os.path.join(LIB, "file_view1")
    if not os.path.exists(exe):
        pytest.skip("lib/file_view1 is not built (the reference's toolchain tests were not at hand)")
    print("[gpu] file_view1", flush=True)
    p = subprocess.run([exe, "---gpumpi", "-n", "4", "-b", "64"], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    data = np.fromfile(str(tmp_path / "test_file_view1.out"), dtype=np.int32)
    assert np.array_equal(data, np.tile(np.arange(100, dtype=np.int32), 4))
