"""Whole rank programs in the CPU simulation of the device runtime
(tests/hostbuild/simt): every lane a fiber, lockstep wave operations, built
under AddressSanitizer and UBSan.  These run the launcher and the programs
end to end without a GPU: the batched uneven Allreduce (exact, BF16 and FP8
wires), the compressed-wire Allreduce, the uneven Alltoallv and the block collectives at the shapes of the
GPU tests, and a launch in which one rank aborts while the others keep
calling collectives."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "hostbuild", "simt")
OUT = os.path.join(HERE, "hostbuild", "_build", "simt")
ENV = dict(os.environ, ASAN_OPTIONS="detect_odr_violation=0:detect_leaks=0", SIMT_TIMEOUT="600")


@pytest.fixture(scope="module")
def sim():
    env = dict(os.environ, SIMT_PROGS="coll_bench.hip batched.hip")
    subprocess.run(["bash", os.path.join(SIM, "build.sh"), OUT, os.path.join(SIM, "run_coll.cpp"),
                    os.path.join(SIM, "run_batched.cpp")], check=True, env=env, timeout=1200)
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


@pytest.mark.parametrize("kind", range(7))
def test_sim_block_collectives(sim, kind):
    rc, out = _run([os.path.join(sim, "run_coll"), "blocks", str(kind), "128"])
    assert rc == 0 and f"blocks kind {kind} mismatches: 0" in out, out[-3000:]


@pytest.mark.parametrize("wire,P,n", [(1, 4, 1000), (2, 4, 1000), (1, 9, 70001), (2, 9, 70001)])
def test_sim_wire_allreduce(sim, wire, P, n):
    # The compressed-wire Allreduce at the GPU test's shapes: every rank's
    # output identical, within the wire's error of the exact sum, and no
    # out-of-bounds access under AddressSanitizer.
    rc, out = _run([os.path.join(sim, "run_coll"), "wire", str(wire), str(P), str(n), "256"])
    assert rc == 0 and f"wire {wire} P={P} n={n}: rank mismatches 0" in out, out[-3000:]


def test_sim_abort_leaves_unpublished_tables_alone(sim):
    # Rank 1 aborts, then reaches MPI_Finalize; its barrier arrival completes
    # the others' Allreduce barrier.  They must see the abort and return
    # instead of folding from table entries rank 1 never published.
    rc, out = _run([os.path.join(sim, "run_coll"), "abort", "8", "128"])
    assert rc == 0, out[-3000:]
    assert "mpi_error=7" in out and "AddressSanitizer" not in out


@pytest.fixture(scope="module")
def sim_is():
    out = os.path.join(HERE, "hostbuild", "_build", "simt_is")
    env = dict(os.environ, SIMT_PROGS="npb_is.hip")
    subprocess.run(["bash", os.path.join(SIM, "build.sh"), out, os.path.join(SIM, "run_is.cpp")], check=True,
                   env=env, timeout=1200)
    return out


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
def test_sim_npb_is_matches_oracle(sim_is, orc, cls, P, ulds):
    res, keys, _ = orc.is_run(cls, P, want_keys=True)
    rc, out = _run([os.path.join(sim_is, "run_is"), cls, str(P), "128", str(ulds)])
    assert rc == 0, out[-3000:]
    assert (f"passed={res.passed_verification} total={keys.size} errors=0 keyhash={_keyhash(keys)}") in out, out[-2000:]
