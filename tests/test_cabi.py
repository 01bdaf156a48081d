"""The C-ABI libraries load and export what include/gpumpi/host.h declares
(no GPU calls), and the host-side configuration logic behaves like the
reference's launcher (gpu_libs/gpu_main/main.cu:48-97)."""
import ctypes
import os
import re

import pytest

import gpumpi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", "gpumpi", header)).read()
    return set(re.findall(r"^\s*(?:int|const char\*|long long|const void\*)\s+\**(gpumpi_\w+)\s*\(", txt, re.M))


@pytest.fixture(scope="module")
def libs():
    if not gpumpi.libs_built():
        pytest.fail("libgpumpi.so / libgpumpi_progs.so are not built (run dphpc-gpu_mpi_amd/build.py)")
    return ctypes.CDLL(gpumpi.CORE_LIB), ctypes.CDLL(gpumpi.PROGS_LIB)


def test_host_header_symbols_exported(libs):
    core, _ = libs
    decl = _declared("host.h")
    assert set(gpumpi.CORE_SYMBOLS) == decl
    for name in decl:
        assert hasattr(core, name), name


def test_progs_symbols_exported(libs):
    _, progs = libs
    for name in gpumpi.PROGS_SYMBOLS:
        assert hasattr(progs, name), name


def test_program_registry(libs):
    names = gpumpi.programs()
    for n in ("ka_single_int", "ka_all_to_all_fifo", "ka_comm_split", "p2p_ring_sweep",
              "p2p_pingpong_sweep", "coll_reduce", "coll_alltoallv", "cpi"):
        assert n in names
    _, progs = gpumpi.load()
    assert progs.gpumpi_prog_lookup(b"no_such_program") is None


def test_default_config():
    c = gpumpi.default_config(256)
    assert c.num_procs == 256 and c.threads_per_rank % 64 == 0
    assert c.chan_slots & (c.chan_slots - 1) == 0
    assert c.heap_bytes > 0 and c.staging_bytes > 0


def _parse(args, nranks=0):
    core = gpumpi.load_core()
    c = gpumpi.default_config(nranks)
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    rc = core.gpumpi_parse_args(len(args), argv, ctypes.byref(c))
    return rc, c


def test_parse_args_like_reference_cli():
    rc, c = _parse(["./pi", "100", "---gpumpi", "-n", "64", "-b", "512", "-e", "300", "-s", "8192"])
    assert rc == 2  # the program's own argc
    assert (c.num_procs, c.threads_per_rank, c.max_requests, c.stack_size) == (64, 512, 300, 8192)
    rc, c = _parse(["./pi", "---gpumpi", "-g", "4"])
    assert rc == 1 and c.num_procs == 4 and c.blocks_per_grid == 4


def test_parse_args_errors():
    rc, _ = _parse(["./pi", "-n", "4"])
    assert rc == -1  # no ---gpumpi marker (main.cu:56-62 exits there)
    assert "---gpumpi" in gpumpi.last_error()
    rc, _ = _parse(["./pi", "---gpumpi", "-n"])
    assert rc == -2
    rc, _ = _parse(["./pi", "---gpumpi", "--bogus", "3"])
    assert rc == -2
