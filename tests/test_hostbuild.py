"""The runtime's lane-local device code built for the host (CPU) and run one
lane at a time (tests/hostbuild/hip/hip_runtime.h):

* the all-to-all / all-gather data phase (coll.hpp seg_info, run_segments,
  copy_part) under AddressSanitizer, for every segment kind, small and large
  segments, so every address the device would touch is bounds-checked;
* copy_part alone over random alignments (its realigning 16-byte path);
* the reduction fold (ops.hpp fold_builtin) for all 216 legal
  (datatype, op) pairs, bit-exact with the oracle's rank-order fold, with
  and without the 16-byte vector path, for Allreduce and rooted Reduce."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HB = os.path.join(HERE, "hostbuild")
OUT = os.path.join(HB, "_build")
CLANG = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-std=c++17", "-g", "-ffp-contract=off", "-I", HB, "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "dphpc-gpu_mpi_amd", "csrc"), "-x", "c++"]


def _build(src, out, extra):
    os.makedirs(OUT, exist_ok=True)
    deps = [os.path.join(HB, src), os.path.join(HB, "hip", "hip_runtime.h")]
    for d, _, fs in os.walk(os.path.join(ROOT, "dphpc-gpu_mpi_amd", "csrc", "gpumpi_rt")):
        deps += [os.path.join(d, f) for f in fs]
    if os.path.exists(out) and all(os.path.getmtime(out) > os.path.getmtime(d) for d in deps):
        return out
    subprocess.run([CLANG] + FLAGS + [os.path.join(HB, src)] + extra + ["-o", out], check=True,
                   capture_output=True, text=True)
    return out


@pytest.fixture(scope="module")
def segcheck():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ not available")
    return _build("segments_check.cpp", os.path.join(OUT, "segments_check"),
                  ["-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"])


@pytest.fixture(scope="module")
def copycheck():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ not available")
    return _build("copy_check.cpp", os.path.join(OUT, "copy_check"),
                  ["-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"])


def test_copy_part_alignments_in_bounds(copycheck):
    """copy_part over random source/destination alignments and lengths: exact
    bytes, nothing written outside the destination, no source read outside
    the aligned 16-byte blocks that hold source bytes (poisoned under ASan)."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([copycheck, "3000"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "mismatches=0" in r.stdout


@pytest.fixture(scope="module")
def hostfold():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ not available")
    so = _build("fold_host.cpp", os.path.join(OUT, "libfold_host.so"),
                ["-O2", "-shared", "-fPIC", "-fsanitize=undefined", "-fsanitize-trap=undefined"])
    lib = ctypes.CDLL(so)
    lib.host_fold.restype = ctypes.c_int
    lib.host_fold.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                              ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
    return lib


# SegKind: 0 ALLTOALLV, 1 ALLTOALL, 2 ALLGATHER, 3 ALLGATHERV
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("P,blk", [(1, 5), (2, 1), (6, 5), (17, 3), (64, 7), (6, 70001)])
def test_segment_copies_in_bounds(segcheck, kind, P, blk):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    r = subprocess.run([segcheck, str(kind), str(P), str(blk)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert "mismatches=0" in r.stdout


def _rand(orc, dtype, P, n, rng):
    npdt = np.dtype(orc.NP_DTYPE[dtype])
    if dtype == "MPI_C_BOOL":
        return rng.integers(0, 2, (P, n)).astype(np.uint8)
    if npdt.kind in "iu":
        info = np.iinfo(npdt)
        return rng.integers(info.min, info.max, (P, n), dtype=npdt, endpoint=True)
    if npdt.kind == "f":
        return (rng.standard_normal((P, n)) * 1e3).astype(npdt)
    return (rng.standard_normal((P, n)) + 1j * rng.standard_normal((P, n))).astype(npdt)


def _aligned(a):
    """Copy of 1-D array `a` at a 16-byte aligned address (the device's
    vector path is taken only for 16-byte aligned buffers)."""
    raw = np.zeros(a.nbytes + 16, dtype=np.uint8)
    off = (-raw.ctypes.data) % 16
    out = raw[off:off + a.nbytes].view(a.dtype)
    out[:] = a
    return out


def _host_fold(lib, orc, dtype, op, x, lo, hi, root, vec, nt=64):
    P = x.shape[0]
    ins = [_aligned(x[r]) for r in range(P)]
    outs = [_aligned(np.zeros_like(x[r])) for r in range(P)]
    ip = (ctypes.c_void_p * P)(*[a.ctypes.data for a in ins])
    op_ = (ctypes.c_void_p * P)(*[a.ctypes.data for a in outs])
    rc = lib.host_fold(orc.DT[dtype], orc.OP[op], P, ip, op_, lo, hi, root, vec, nt)
    assert rc == 0
    return outs


def test_fold_every_legal_pair_matches_oracle(hostfold, orc):
    rng = np.random.default_rng(23)
    P, n = 5, 37
    pairs = [(t, o) for t in orc.DTYPES for o in orc.OPS if orc.op_allowed(t, o)]
    assert len(pairs) == 216
    for dtype, op in pairs:
        x = _rand(orc, dtype, P, n, rng)
        exp = orc.reduce_fold(dtype, op, x)
        for vec in (0, 1):
            outs = _host_fold(hostfold, orc, dtype, op, x, 0, n, -1, vec)
            for r in range(P):
                assert outs[r].tobytes() == exp.tobytes(), (dtype, op, vec, r)
        outs = _host_fold(hostfold, orc, dtype, op, x, 0, n, 2, 1)
        assert outs[2].tobytes() == exp.tobytes(), (dtype, op, "root")
        assert not any(outs[r].any() for r in (0, 1, 3, 4))


@pytest.mark.parametrize("dtype,op", [("MPI_DOUBLE", "MPI_SUM"), ("MPI_FLOAT", "MPI_PROD"), ("MPI_INT8_T", "MPI_BXOR"),
                                      ("MPI_C_DOUBLE_COMPLEX", "MPI_PROD"), ("MPI_SHORT", "MPI_MIN")])
def test_fold_member_slices_cover_the_vector(hostfold, orc, dtype, op):
    """Members fold disjoint 16-byte-group slices (reduce_common's split);
    together they produce the whole result, for several worker counts."""
    rng = np.random.default_rng(5)
    P, n = 7, 1029
    x = _rand(orc, dtype, P, n, rng)
    exp = orc.reduce_fold(dtype, op, x)
    es = x.dtype.itemsize
    V = 1 if es >= 16 else 16 // es
    groups = (n + V - 1) // V
    for nt in (64, 256):
        res = np.zeros_like(x[0])
        for me in range(P):
            lo, hi = groups * me // P * V, min(n, groups * (me + 1) // P * V)
            if lo >= hi:
                continue
            outs = _host_fold(hostfold, orc, dtype, op, x, lo, hi, -1, 1, nt)
            res[lo:hi] = outs[0][lo:hi]
            for r in range(P):   # nothing outside the slice is written
                assert not outs[r][:lo].any() and not outs[r][hi:].any()
        assert res.tobytes() == exp.tobytes(), nt
