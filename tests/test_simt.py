"""Whole rank programs in the CPU simulation of the device runtime
(tests/hostbuild/simt): every lane a fiber, lockstep wave operations, built
under AddressSanitizer and UBSan.  These run the launcher and the programs
end to end without a GPU: the batched uneven Allreduce (exact, BF16 and FP8
wires), the compressed-wire Allreduce, the uneven Alltoallv and the block collectives at the shapes of the
GPU tests, and a launch in which one rank aborts while the others keep
calling collectives.  With SIMT_RACE=1 the simulation also checks every
cross-rank access for happens-before ordering (simt.cpp)."""
import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "hostbuild", "simt")
OUT = os.path.join(HERE, "hostbuild", "_build", "simt")
ENV = dict(os.environ, ASAN_OPTIONS="detect_odr_violation=0:detect_leaks=0", SIMT_TIMEOUT="600")


REF_CPI = "/root/reference/toolchain_tests/pi/cpi.c"


SIM_MAX_CTX = 255
REF_TOOLCHAIN = "/root/reference/toolchain_tests"
NPB = "/root/reference/toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI"


def _pkg_build():
    spec = importlib.util.spec_from_file_location("gpumpi_build", os.path.join(HERE, "..", "dphpc-gpu_mpi_amd",
                                                                               "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _convert(converter, src, dst, incs=()):
    """converter.convert into dst, leaving dst (and its timestamp) alone when
    the text is the same, so that the incremental simulation build keeps its
    object."""
    tmp = dst + ".new"
    converter.convert(src, tmp, list(incs))
    if os.path.exists(dst) and open(dst).read() == open(tmp).read():
        os.remove(tmp)
    else:
        os.replace(tmp, dst)


@pytest.fixture(scope="module")
def sim():
    # One build of every simulated program and driver: build.sh compiles all
    # translation units in parallel.  The drivers include C programs run
    # through the source converter (each has its own host main()).
    import converter
    os.makedirs(OUT, exist_ok=True)
    drivers = [os.path.join(SIM, d) for d in ("run_coll.cpp", "run_batched.cpp", "run_is.cpp", "run_io.cpp",
                                              "run_pic.cpp", "run_lifetime.cpp", "run_user_op.cpp",
                                              "run_p2p.cpp", "run_pic_pdecomp.cpp", "run_semantics.cpp",
                                              "run_known_answer.cpp")]
    ring = os.path.join(OUT, "conv_ring.cpp")
    _convert(converter, os.path.join(HERE, "..", "dphpc-gpu_mpi_amd", "examples", "ring.c"), ring)
    drivers.append(ring)
    if os.path.exists(REF_CPI):
        cpi = os.path.join(OUT, "conv_cpi.cpp")
        _convert(converter, REF_CPI, cpi)
        drivers.append(cpi)
    # Context ids limited to 3..255 here, so that id reuse is reached in a few
    # hundred communicator creations.
    for rel, name in (("global_var/global_var.c", "conv_global_var.cpp"), ("basicIO/basicIO.c", "conv_basicio.cpp"),
                      ("datatype/contiguous.c", "conv_contiguous.cpp"),
                      ("basicIO/file_view1.c", "conv_file_view1.cpp")):
        src = os.path.join(REF_TOOLCHAIN, rel)
        if os.path.exists(src):
            _convert(converter, src, os.path.join(OUT, name))
            drivers.append(os.path.join(OUT, name))
    derived = os.path.join(OUT, "conv_derived.cpp")
    _convert(converter, os.path.join(HERE, "programs", "derived_types.c"), derived)
    drivers.append(derived)
    exit_prog = os.path.join(OUT, "conv_exit.cpp")
    _convert(converter, os.path.join(HERE, "programs", "exit_early.c"), exit_prog)
    drivers.append(exit_prog)
    cflags = f"-DGPUMPI_MAX_CTX={SIM_MAX_CTX}"
    if os.path.exists(os.path.join(NPB, "DT", "dt.c")):
        # the reference's NPB DT, unmodified, as one converted program
        inc = _pkg_build().npb_params("dt", "S")
        parts = []
        for src in ("DT/dt.c", "DT/DGraph.c", "common/c_print_results.c", "common/c_timers.c", "common/randdp.c"):
            dst = os.path.join(OUT, "conv_" + os.path.basename(src).replace(".c", ".cpp"))
            _convert(converter, os.path.join(NPB, src), dst, [inc])
            parts.append(dst)
        drivers.append(":".join(parts))
        cflags += f" -I {NPB}/DT -I {NPB}/common -I {inc}"
    env = dict(os.environ, SIMT_PROGS="coll_bench.hip batched.hip npb_is.hip io.hip pic.hip pic_exact.hip lifetime.hip "
               "user_ops.hip p2p_bench.hip pic_pdecomp.hip semantics.hip known_answer.hip", SIMT_CFLAGS=cflags)
    subprocess.run(["bash", os.path.join(SIM, "build.sh"), OUT] + drivers, check=True, env=env, timeout=1200)
    return OUT


def _run(args, timeout=900):
    p = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def _batched_layout(P, seed, nmin_log2, nmax_log2):
    # same generator as tests/test_gpu_wire.py
    rng = np.random.default_rng(seed)
    sizes = []
    while sum(sizes) < P:
        sizes.append(int(rng.integers(4, 65)))
    excess = sum(sizes) - P
    sizes[-1] -= excess
    if sizes[-1] < 4:
        sizes[-2] += sizes[-1]
        sizes.pop()
    colors = np.repeat(np.arange(len(sizes)), sizes).astype(np.int32)
    counts = np.floor(2.0 ** rng.uniform(nmin_log2, nmax_log2, len(sizes))).astype(np.int64)
    offs, o = [], 0
    for r in range(P):
        offs.append(o)
        o += int(counts[colors[r]] + 63) // 64 * 64
    return colors, counts, offs, o


def test_sim_batched_allreduce_all_wires(sim, tmp_path):
    colors, counts, offs, total = _batched_layout(40, 3, 8, 13)
    f = tmp_path / "layout.txt"
    f.write_text(f"{len(colors)} {len(counts)} {total}\n" + " ".join(map(str, colors)) + "\n" +
                 " ".join(map(str, counts)) + "\n" + " ".join(map(str, offs)) + "\n")
    rc, out = _run([os.path.join(sim, "run_batched"), str(f), "128", "64", str(4 << 20), "2", "0", "1", "2"])
    assert rc == 0, out[-3000:]
    assert "wire=0 output mismatches vs rank-order fold: 0" in out
    assert out.count(" rc=0 status=0 mpi_error=0") == 3, out[-3000:]


@pytest.mark.parametrize("P", [4, 16])
def test_sim_alltoallv_uneven(sim, tmp_path, P):
    rng = np.random.default_rng(P)  # the count matrix of tests/test_gpu_coll.py
    counts = rng.integers(0, 3000, (P, P)).astype(np.int32)
    counts[rng.random((P, P)) < 0.1] = 0
    f = tmp_path / "counts.txt"
    f.write_text(f"{P}\n" + " ".join(map(str, counts.ravel())) + "\n")
    rc, out = _run([os.path.join(sim, "run_coll"), "alltoallv", str(f), "256"])
    assert rc == 0 and "alltoallv mismatches: 0" in out, out[-3000:]


@pytest.mark.parametrize("kind", range(9))
def test_sim_block_collectives(sim, kind):
    # kinds 7 and 8: Alltoall and Alltoallv with MPI_IN_PLACE
    rc, out = _run([os.path.join(sim, "run_coll"), "blocks", str(kind), "128"])
    assert rc == 0 and f"blocks kind {kind} mismatches: 0" in out, out[-3000:]


@pytest.mark.parametrize("wire,P,n", [(1, 4, 1000), (2, 4, 1000), (1, 9, 70001), (2, 9, 70001)])
def test_sim_wire_allreduce(sim, wire, P, n):
    # The compressed-wire Allreduce at the GPU test's shapes: every rank's
    # output identical, within the wire's error of the exact sum, and no
    # out-of-bounds access under AddressSanitizer.
    rc, out = _run([os.path.join(sim, "run_coll"), "wire", str(wire), str(P), str(n), "256"])
    assert rc == 0 and f"wire {wire} P={P} n={n}: rank mismatches 0" in out, out[-3000:]


def _run_race(args, timeout=900):
    env = dict(ENV, SIMT_RACE="1")
    p = subprocess.run(args, capture_output=True, text=True, env=env, timeout=timeout)
    return p.returncode, p.stdout + p.stderr


def test_sim_race_checker_flags_a_planted_race(sim):
    # The happens-before checker itself: a write by rank 0 that rank 1 reads
    # with no synchronisation between them is reported with both sites; the
    # same accesses separated by MPI_Barrier are not.
    rc, out = _run_race([os.path.join(sim, "run_coll"), "race_selftest", "0"])
    assert rc == 0, out[-3000:]
    assert "simt race: read of an unordered write" in out and "run_coll.cpp" in out, out[-3000:]
    assert "simt race: write after an unordered read" in out
    rc, out = _run_race([os.path.join(sim, "run_coll"), "race_selftest", "1"])
    assert rc == 0 and "simt race:" not in out and "0 report(s)" in out, out[-3000:]


@pytest.mark.parametrize("case", [["blocks", str(k), "128"] for k in range(7)] +
                         [["wire", "1", "4", "1000", "256"], ["wire", "2", "9", "2000", "256"]],
                         ids=lambda c: "-".join(c))
def test_sim_collectives_have_no_unordered_accesses(sim, case):
    # Every plain access a rank makes to another rank's data (table words,
    # member buffers, the communicator's shared state) is ordered by the
    # agent-scope release/acquire of a barrier or a channel word.
    rc, out = _run_race([os.path.join(sim, "run_coll")] + case)
    assert rc == 0, out[-3000:]
    assert "simt race:" not in out and "race check: 0 report(s)" in out, out[-3000:]


def test_sim_alltoallv_has_no_unordered_accesses(sim, tmp_path):
    P = 4
    rng = np.random.default_rng(P)
    counts = rng.integers(0, 3000, (P, P)).astype(np.int32)
    counts[rng.random((P, P)) < 0.1] = 0
    f = tmp_path / "counts.txt"
    f.write_text(f"{P}\n" + " ".join(map(str, counts.ravel())) + "\n")
    rc, out = _run_race([os.path.join(sim, "run_coll"), "alltoallv", str(f), "256"])
    assert rc == 0 and "alltoallv mismatches: 0" in out, out[-3000:]
    assert "simt race:" not in out and "race check: 0 report(s)" in out, out[-3000:]


def test_sim_abort_leaves_unpublished_tables_alone(sim):
    # Rank 1 aborts, then reaches MPI_Finalize; its barrier arrival completes
    # the others' Allreduce barrier.  They must see the abort and return
    # instead of folding from table entries rank 1 never published.
    # Under the happens-before checker: no rank reads what the aborting
    # rank never published, and nothing it did publish unordered.
    rc, out = _run_race([os.path.join(sim, "run_coll"), "abort", "8", "128"])
    assert rc == 0, out[-3000:]
    assert "mpi_error=7" in out and "AddressSanitizer" not in out
    assert "simt race:" not in out and "race check: 0 report(s)" in out, out[-3000:]


def _keyhash(keys):
    h = 1469598103934665603
    for x in keys.astype(np.uint32).tolist():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


# (class, ranks, user LDS bytes): 4096 keeps the counts and the shared
# bucket histogram in global memory / one LDS array; the larger sizes put
# per-wave sub-histograms and (where the key range fits) the ranking counts
# in LDS.
@pytest.mark.parametrize("cls,P,ulds", [("S", 4, 4096), ("S", 4, 20480), ("S", 16, 20480), ("W", 4, 36864)])
def test_sim_npb_is_matches_oracle(sim, orc, cls, P, ulds):
    res, keys, _ = orc.is_run(cls, P, want_keys=True)
    rc, out = _run([os.path.join(sim, "run_is"), cls, str(P), "128", str(ulds)])
    assert rc == 0, out[-3000:]
    assert (f"passed={res.passed_verification} total={keys.size} errors=0 keyhash={_keyhash(keys)}") in out, out[-2000:]


@pytest.mark.parametrize("P,N", [(1, 5), (4, 10), (9, 300)])
def test_sim_io_scenarios(sim, tmp_path, P, N):
    # MPI_File_* and the printf family on the host service channel: the
    # reference's IO test cases plus explicit offsets, the shared pointer,
    # multi-slot transfers, append/exclusive opens and formatted text.
    import io_expect
    rc, out = _run([os.path.join(sim, "run_io"), str(P), str(tmp_path), str(N)])
    assert rc == 0, out[-3000:]
    recs, texts = {}, {}
    for line in out.splitlines():
        if line.startswith("rank "):
            r, vals = line[5:].split(":", 1)
            recs[int(r)] = [int(v) for v in vals.split()]
        elif line.startswith("text "):
            r, t = line[5:].split(": ", 1)
            texts[int(r)] = t
    io_expect.check([recs[r] for r in range(P)], [texts[r] for r in range(P)], P, N)
    for r in range(P):
        assert f"io_scenarios: rank {r} of {P}" in out
    assert os.listdir(tmp_path) == []  # every file the program made is deleted


# The converted programs keep their data on the rank heap and in per-rank
# globals, which every lane of the program's wave updates together; a heap
# below the simulation's bulk-allocation size keeps those accesses in lane
# lockstep, as the hardware runs them.
SMALL_HEAP = ["-p", str(4 << 20)]


@pytest.mark.parametrize("P", [1, 4, 7])
def test_sim_converted_c_program_matches_host_mpi(sim, P):
    # dphpc-gpu_mpi_amd/examples/ring.c through converter.py: per-rank
    # globals, a static local, malloc/calloc/memset, a C++-keyword variable,
    # MPI_Sendrecv around a ring and MPI_Allreduce; the output line must equal
    # the same C program's under host MPICH (tests/golden/ring_mpich.json).
    import json
    gold = json.load(open(os.path.join(HERE, "golden", "ring_mpich.json")))
    rc, out = _run([os.path.join(sim, "conv_ring"), "5", "---gpumpi", "-n", str(P), "-b", "64"] + SMALL_HEAP)
    assert rc == 0, out[-3000:]
    assert gold[str(P)] in out, (gold[str(P)], out[-2000:])


@pytest.mark.skipif(not os.path.exists(REF_CPI), reason="the reference's cpi.c is not at hand")
def test_sim_reference_cpi_converts_and_runs(sim, orc):
    # The reference's own toolchain_tests/pi/cpi.c, unmodified, through the
    # converter: every rank reports in, and pi is the rank-order FP64 fold of
    # the cyclic partials (the oracle, which host MPICH's cpi pins).
    import re
    P, n = 4, 100000
    rc, out = _run([os.path.join(sim, "conv_cpi"), str(n), "---gpumpi", "-n", str(P), "-b", "64"] + SMALL_HEAP)
    assert rc == 0, out[-3000:]
    for r in range(P):
        assert f"Process {r} of {P} is on" in out
    pi = float(re.search(r"pi is approximately ([0-9.]+)", out).group(1))
    want, _ = orc.pi_fold(n, P)
    assert pi == float(f"{want:.16f}"), (pi, want)


def _pic_out(out):
    e, e16, ranks = {}, {}, {}
    for line in out.splitlines():
        f = line.split()
        if f[:1] == ["energy"]:
            e[int(f[1])] = int(f[2])
        elif f[:1] == ["energy16"]:
            e16[int(f[1])] = int(f[2])
        elif f[:1] == ["rank"] and "phash" in f:
            ranks[int(f[1])] = (int(f[3]), int(f[5]), int(f[7]))
    return e, e16, ranks


@pytest.mark.parametrize("shape", [(2, 2, 16, 16, 2000, 5), (3, 2, 8, 12, 500, 12), (2, 1, 12, 8, 400, 8),
                                   (1, 1, 10, 10, 300, 4)])
def test_sim_pic_matches_oracle(sim, orc, shape):
    # The C5 rank program (progs/pic.hip) against orc_pic_run, bit for bit:
    # every cycle's FP64 energy and BF16-wire energy, and per rank the final
    # particle count, particle array and density (with its halo ring).
    _check_pic(sim, orc, shape, "pic")


@pytest.mark.parametrize("shape", [(2, 2, 16, 16, 2000, 5), (3, 3, 8, 8, 300, 3)])
def test_sim_pic_exact_matches_oracle(sim, orc, shape):
    # `pic_exact` (the same cycle without the BF16-wire Allreduce) leaves the
    # BF16 energies unwritten and matches the oracle in everything else.
    _check_pic(sim, orc, shape, "pic_exact")


def _check_pic(sim, orc, shape, prog):
    import struct
    px, py, nx, ny, np_, cycles = shape
    rc, out = _run([os.path.join(sim, "run_pic"), str(px), str(py), str(nx), str(ny), str(np_), str(cycles),
                    "2024", "128", prog])
    assert rc == 0, out[-3000:]
    e, e16, ranks = _pic_out(out)
    d = orc.pic_run(orc.pic_params(px, py, nx, ny, np_, cycles, seed=2024))
    for c in range(cycles):
        assert e[c] == struct.unpack("<Q", struct.pack("<d", d["energy"][c]))[0], c
        want16 = struct.unpack("<I", struct.pack("<f", d["energy_bf16"][c]))[0] if prog == "pic" else 0
        assert e16[c] == want16, c
    for r in range(px * py):
        assert ranks[r] == (int(d["counts"][r]), int(d["part_hash"][r]), int(d["rho_hash"][r])), r
    tp = [float(x) for x in next(l for l in out.splitlines() if l.startswith("phases")).split()[1:]]
    assert len(tp) == 7 and tp[6] > 0 and min(tp) >= 0 and sum(tp[:6]) <= tp[6] * 1.0001, tp


def test_sim_pic_has_no_unordered_accesses(sim):
    rc, out = _run_race([os.path.join(sim, "run_pic"), "3", "3", "8", "8", "300", "3", "7", "128"])
    assert rc == 0, out[-3000:]
    assert "simt race:" not in out and "race check: 0 report(s)" in out, out[-3000:]


def _lifetime(sim, mode, P, iters, pool=None, race=False):
    cmd = [os.path.join(sim, "run_lifetime"), str(mode), str(P), str(iters), "64"] + ([str(pool)] if pool else [])
    rc, out = (_run_race if race else _run)(cmd)
    assert rc == 0, out[-3000:]
    oks = [int(l.split()[2]) for l in out.splitlines() if l.startswith("ok ")]
    outs = [int(x) for x in next(l for l in out.splitlines() if l.startswith("out ")).split()[1:]]
    return oks, outs, out


@pytest.mark.parametrize("P", [1, 3])
def test_sim_rank_heap_reuses_merges_and_lowers_the_top(sim, P):
    # Exact block positions after frees, splits, merges and a lowered top;
    # then 2500 random allocations and frees (up to 16 blocks of up to 64 KiB
    # live, each painted and checked) in a 2 MiB heap, which a bump
    # allocator would exhaust; the heap is empty again at the end.
    oks, outs, _ = _lifetime(sim, 0, P, 2500)
    assert oks == [1] * P and outs[0] == 2500 and outs[1] >= 8


@pytest.mark.parametrize("P", [2, 5])
def test_sim_comm_free_returns_blocks_and_context_ids(sim, P):
    # More communicator creations than there are context ids (3..255 in this
    # build) and than a 64 KiB pool holds blocks for: every one is freed.
    iters = SIM_MAX_CTX + 40 if P == 2 else 60
    oks, outs, _ = _lifetime(sim, 1, P, iters, pool=64 << 10)
    assert oks == [1] * P and outs[0] == iters


def test_sim_comm_free_completes_pending_requests(sim):
    oks, _, out = _lifetime(sim, 2, 4, 0, race=True)
    assert oks == [1] * 4
    assert "simt race:" not in out and "race check: 0 report(s)" in out, out[-3000:]


def _user_op_expected_sum(P, n):
    # x_0 + 2 * (x_1 + ... + x_{P-1}) per element, summed: the reference's
    # fold (acc = x_0, then fn(x_r, acc) for r = 1 .. P-1) with
    # fn: inout = 2 * in + inout.  The reversed fold would weight x_{P-1} once.
    i = np.arange(n, dtype=np.int64)
    x = lambda r: (r * 7919 + i * 104729) % 100003 - 50000  # noqa: E731
    return int((x(0) + 2 * sum((x(r) for r in range(1, P)), np.zeros(n, dtype=np.int64))).sum())


@pytest.mark.parametrize("P,n,root,in_place", [(1, 10, -1, 1), (2, 1, 1, 0), (3, 1000, -1, 0), (3, 1000, -1, 1),
                                               (4, 63, 2, 1), (5, 4097, 0, 0), (6, 300, 5, 0)])
def test_sim_user_defined_operators(sim, P, n, root, in_place):
    # In the program: the user FP64 sum equals MPI_SUM bit for bit at every
    # receiving rank, and the asymmetric int64 operator equals its closed form.
    rc, out = _run([os.path.join(sim, "run_user_op"), str(P), str(n), str(root), str(in_place), "64"])
    assert rc == 0, out[-3000:]
    oks = [int(l.split()[2]) for l in out.splitlines() if l.startswith("ok ")]
    assert oks == [1] * P, out[-2000:]
    got = int(next(l for l in out.splitlines() if l.startswith("sum ")).split()[1])
    assert got == _user_op_expected_sum(P, n)


@pytest.mark.parametrize("case", [("ring", 4, 100000, 256, 448, 16384, 8), ("ring", 3, 5000, 128, 192, 4096, 4),
                                  ("pingpong", 4, 3000, 64, 4032, 262144, 16), ("ring", 5, 70000, 192, 1984, 65536, 4),
                                  ("pingpong", 2, 1000, 512, 960, 4096, 16), ("ring", 4, 200, 64, 192, 4096, 4),
                                  ("ring", 3, 4032, 128, 4032, 65536, 16)])
def test_sim_p2p_under_every_tuner_candidate(sim, case):
    # The autotuner (gpumpi/tune.py) launches the ring under each candidate
    # of threads per rank, copy-split threshold, eager threshold and channel
    # slots; every candidate value here moves the payloads intact, on both
    # sides of the eager threshold.
    prog, P, nbytes, threads, eager, helper, slots = case
    rc, out = _run([os.path.join(sim, "run_p2p"), prog, str(P), str(nbytes), str(threads), str(eager), str(helper),
                    str(slots)])
    assert rc == 0, out[-3000:]
    assert "mismatches 0" in out, out[-2000:]


NPB_IS = NPB


@pytest.mark.skipif(not os.environ.get("GPUMPI_SIM_SLOW"),
                    reason="about 6 minutes: set GPUMPI_SIM_SLOW=1 to run the converted NPB IS in the simulation")
@pytest.mark.skipif(not os.path.exists(os.path.join(NPB_IS, "IS", "is.c")), reason="the reference's NPB IS is absent")
def test_sim_reference_npb_is_converts_and_verifies(tmp_path):
    # The reference's is.c, c_print_results.c and c_timers.c, unmodified,
    # through the converter; class S on four ranks passes NPB's verification
    # (full_verify receives into a stack variable with MPI_Irecv).
    import converter
    inc = _pkg_build().npb_params("is", "S")
    parts = []
    for src in ("IS/is.c", "common/c_print_results.c", "common/c_timers.c"):
        dst = str(tmp_path / ("conv_" + os.path.basename(src).replace(".c", ".cpp")))
        converter.convert(os.path.join(NPB_IS, src), dst, [inc])
        parts.append(dst)
    env = dict(os.environ, SIMT_PROGS="io.hip",
               SIMT_CFLAGS=f"-I {NPB_IS}/IS -I {NPB_IS}/common -I {inc}")
    out = str(tmp_path / "out")
    subprocess.run(["bash", os.path.join(SIM, "build.sh"), out, ":".join(parts)], check=True, env=env, timeout=1200)
    rc, text = _run([os.path.join(out, "conv_is"), "---gpumpi", "-n", "4", "-b", "64", "-p", str(4 << 20)],
                    timeout=1200)
    assert rc == 0, text[-3000:]
    assert "Verification    =               SUCCESSFUL" in text, text[-3000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(NPB, "DT", "dt.c")), reason="the reference's NPB DT is absent")
@pytest.mark.parametrize("graph", ["BH", "WH", "SH"])
def test_sim_reference_npb_dt_matches_host_mpi(sim, graph):
    # The reference's NPB DT class S (dt.c, DGraph.c and the common files,
    # unmodified) through the converter: the L2 norm host MPICH prints for
    # the same graph, and NPB's verification.
    import json
    gold = json.load(open(os.path.join(HERE, "golden", "npb_dt_mpich.json")))["runs"][graph]
    rc, out = _run([os.path.join(sim, "conv_dt"), graph, "---gpumpi", "-n", str(gold["ranks"]), "-b", "64"]
                   + SMALL_HEAP, timeout=1200)
    assert rc == 0, out[-3000:]
    assert f"DT_{graph}.S L2 Norm = {gold['l2_norm']}" in out, out[-3000:]
    assert "Verification    =               SUCCESSFUL" in out


def _fnv1a(data):
    h = 1469598103934665603
    for x in data:
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("shape,P", [((2, 2, 16, 16, 500, 5), 3), ((3, 2, 8, 12, 100, 7), 4),
                                     ((1, 1, 10, 10, 64, 4), 5)])
def test_sim_particle_decomposition_pic_matches_the_global_oracle(sim, orc, shape, P):
    # sputniPIC's pattern (field Bcast, particle Scatterv / Gatherv with uneven
    # counts, FP32 Reduce of the counts, FP64 Reduce of the energy) against
    # orc_pic_global: the final particles bit for bit, in order; the energy
    # within FP64 summation order.
    import struct
    rc, out = _run([os.path.join(sim, "run_pic_pdecomp")] + [str(v) for v in shape] + ["2024", str(P), "64"])
    assert rc == 0, out[-3000:]
    e = [struct.unpack("<d", struct.pack("<Q", int(l.split()[2])))[0] for l in out.splitlines()
         if l.startswith("energy ")]
    h = int(next(l for l in out.splitlines() if l.startswith("particles ")).split()[3])
    eg, qg = orc.pic_global(orc.pic_params(*shape, seed=2024))
    assert h == _fnv1a(np.asarray(qg, dtype=np.float32).tobytes())
    assert len(e) == shape[5] and max(abs(a - b) / b for a, b in zip(e, eg)) <= 1e-13


def test_sim_converted_exit_ends_only_the_calling_rank(sim):
    # exit() in a converted program ends the calling rank's program with that
    # value (its exit work, then s_endpgm): the rest of that rank's main does
    # not run, the other ranks finish, and the executable reports a rank
    # that returned non-zero.
    rc, out = _run([os.path.join(sim, "conv_exit"), "---gpumpi", "-n", "4", "-b", "64"] + SMALL_HEAP)
    assert rc == 1, out[-3000:]
    assert "exit_early: size 4 sum 6" in out and out.count("rank leaving with 3") == 1
    printed = sorted(int(l.split(": ")[1]) for l in out.splitlines() if l.startswith("never printed by rank 1"))
    assert printed == [0, 2, 3], out[-3000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_TOOLCHAIN, "global_var", "global_var.c")),
                    reason="the reference's toolchain tests are absent")
def test_sim_reference_global_var_test_passes(sim):
    # The reference's own converter test (toolchain_tests/global_var, run by
    # its test.sh with 4 ranks): per-rank globals, a global written through a
    # macro, several declarators in one declaration, static locals; it
    # returns non-zero and prints "error ..." on any mismatch.
    rc, out = _run([os.path.join(sim, "conv_global_var"), "---gpumpi", "-n", "4", "-b", "64"] + SMALL_HEAP)
    assert rc == 0 and "error" not in out.replace("runtime error", ""), out[-3000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_TOOLCHAIN, "basicIO", "basicIO.c")),
                    reason="the reference's toolchain tests are absent")
def test_sim_reference_basic_io_test_writes_its_file(sim, tmp_path):
    # toolchain_tests/basicIO/basicIO.c: rank 0 writes its greeting into
    # test.out through MPI_File_write on a file every rank opened.
    p = subprocess.run([os.path.join(sim, "conv_basicio"), "---gpumpi", "-n", "4", "-b", "64"] + SMALL_HEAP,
                       capture_output=True, text=True, env=ENV, timeout=900, cwd=str(tmp_path))
    assert p.returncode == 0, (p.stdout + p.stderr)[-3000:]
    assert (tmp_path / "test.out").read_bytes() == b"hello world from GPUMPI IO"


@pytest.mark.parametrize("P", [2, 3, 4])
def test_sim_converted_contiguous_datatypes(sim, P):
    # tests/programs/derived_types.c (passes under host MPICH at 3 and 4
    # ranks): contiguous types in sizes, nesting, Send/Recv into a stack
    # buffer, Get_count, Bcast, and Allreduce/Reduce with user operators.
    rc, out = _run([os.path.join(sim, "conv_derived"), "---gpumpi", "-n", str(P), "-b", "64"] + SMALL_HEAP)
    assert rc == 0 and "derived_types: ok" in out, out[-3000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_TOOLCHAIN, "datatype", "contiguous.c")),
                    reason="the reference's toolchain tests are absent")
def test_sim_reference_contiguous_test(sim):
    # toolchain_tests/datatype/contiguous.c: one MPI_Type_contiguous(100,
    # MPI_CHAR) element from rank 0 to rank 1 (the reference left the type
    # constructor NOT_IMPLEMENTED).
    rc, out = _run([os.path.join(sim, "conv_contiguous"), "---gpumpi", "-n", "2", "-b", "64"] + SMALL_HEAP)
    assert rc == 0, out[-3000:]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_TOOLCHAIN, "basicIO", "file_view1.c")),
                    reason="the reference's toolchain tests are absent")
def test_sim_reference_file_view_test(sim, tmp_path):
    # toolchain_tests/basicIO/file_view1.c: each rank's 100 ints 0..99 at
    # byte rank*400 through a view whose filetype is a contiguous type.
    p = subprocess.run([os.path.join(sim, "conv_file_view1"), "---gpumpi", "-n", "4", "-b", "64"] + SMALL_HEAP,
                       capture_output=True, text=True, env=ENV, timeout=900, cwd=str(tmp_path))
    assert p.returncode == 0, (p.stdout + p.stderr)[-3000:]
    data = np.fromfile(str(tmp_path / "test_file_view1.out"), dtype=np.int32)
    assert np.array_equal(data, np.tile(np.arange(100, dtype=np.int32), 4))


# The reference's own known-answer tests (tests/cuda_mpi_test.cu,
# collectives_test.cu, communicator_test.cu), the same cases and rank counts
# as tests/test_gpu_known_answer.py.
KNOWN_ANSWER = [("ka_single_int", 2, None), ("ka_array16", 2, None), ("ka_sendrecv_pm", 2, None),
                ("ka_repeat_tags", 2, None), ("ka_network_flood", 3, None), ("ka_all_to_all_fifo", 10, 400),
                ("ka_reduce_prod_i16", 5, None), ("ka_allreduce_double", 7, None), ("ka_allgather", 7, None),
                ("ka_scatter", 5, None), ("ka_scatterv", 5, None), ("ka_comm_rank", 5, None),
                ("ka_comm_dup_isolation", 2, None), ("ka_comm_split", 10, None)]


@pytest.mark.parametrize("name,P,max_requests", KNOWN_ANSWER, ids=[c[0] for c in KNOWN_ANSWER])
def test_sim_reference_known_answer_tests(sim, name, P, max_requests):
    args = [os.path.join(sim, "run_known_answer"), name, str(P), "64"]
    if max_requests:
        args.append(str(max_requests))
    rc, out = _run(args)
    assert rc == 0 and "status=0" in out, out[-3000:]
    oks = [int(l.split()[2]) for l in out.splitlines() if l.startswith("ok ")]
    assert oks == [1] * P, out[-2000:]


SEM_GOLDEN = os.path.join(HERE, "golden", "semantics_mpich.json")


def _sim_semantics(sim, P, threads, mask, race=False):
    rc, out = (_run_race if race else _run)([os.path.join(sim, "run_semantics"), str(P), str(threads), str(mask)])
    assert rc == 0 and "status=0" in out, out[-3000:]
    recs = {}
    for line in out.splitlines():
        if line.startswith("rank ") and ":" in line:
            head, _, vals = line.partition(":")
            recs[int(head.split()[1])] = [int(v) for v in vals.split()]
    return [recs[r] for r in range(P)], out


@pytest.mark.parametrize("P,threads", [(4, 64), (8, 256), (13, 64)])
def test_sim_semantics_match_host_mpi(sim, P, threads):
    # Every rank's record of the p2p / request / communicator scenarios equals
    # host MPICH's (the fixture of tests/test_gpu_semantics.py).
    import json
    gold = json.load(open(SEM_GOLDEN))["runs"][str(P)]
    got, out = _sim_semantics(sim, P, threads, gold["mask"])
    for r in range(P):
        assert got[r] == gold["records"][r], f"rank {r}: {got[r]} != {gold['records'][r]}"


def test_sim_semantics_have_no_unordered_accesses(sim):
    import json
    gold = json.load(open(SEM_GOLDEN))["runs"]["8"]
    got, out = _sim_semantics(sim, 8, 64, gold["mask"], race=True)
    assert got == gold["records"]
    assert "simt race:" not in out and "0 report(s)" in out, out[-3000:]


@pytest.mark.parametrize("rendezvous", [0, 1])
def test_sim_truncated_receive_is_reported(sim, rendezvous):
    # A receive buffer smaller than the message (eager and rendezvous sizes):
    # the launch ends with MPI_ERR_TRUNCATE recorded for the receiving rank,
    # as tests/test_gpu_semantics.py asserts on the GPU.
    rc, out = _run([os.path.join(sim, "run_semantics"), "truncate", str(rendezvous)])
    assert rc == 0, out[-3000:]
    assert "status=4 mpi_error=20 mpi_error_rank=0" in out, out[-3000:]


@pytest.mark.parametrize("threads", [64, 128])
def test_sim_alltoall_truncation_is_an_error(sim, threads):
    # An Alltoall whose send block exceeds the receive block ends the launch
    # with MPI_ERR_TRUNCATE instead of cutting the block short (the case of
    # tests/test_gpu_coll.py::test_alltoall_truncation_is_an_error).
    rc, out = _run([os.path.join(sim, "run_coll"), "truncate", str(threads)])
    assert rc == 0, out[-3000:]


@pytest.mark.parametrize("name,P", [("ka_network_flood", 3), ("ka_comm_split", 10), ("ka_scatterv", 5),
                                    ("ka_comm_dup_isolation", 2), ("ka_repeat_tags", 2)])
def test_sim_known_answer_tests_have_no_unordered_accesses(sim, name, P):
    rc, out = _run_race([os.path.join(sim, "run_known_answer"), name, str(P), "64"])
    assert rc == 0 and "status=0" in out, out[-3000:]
    assert "simt race:" not in out and "0 report(s)" in out, out[-3000:]
