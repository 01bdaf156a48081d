"""Pins the CPU oracle before it is trusted as the checker.

Each check is independent of the oracle's own code: golden vectors produced
by host MPI (MPICH) from the reference's programs or ours, the reference's
known-answer arrays, closed forms, and library routines.
"""
import ctypes
import json
import math
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


# --------------------------------------------------------------------------
# Reductions: every legal (datatype, op) pair against MPICH's own MPI_Reduce
# on the same inputs (tests/golden/reduce_mpich.bin, oracle/hostmpi/
# reduce_golden.c), 5 ranks x 14 elements.
# --------------------------------------------------------------------------
def _golden_inputs(orc, t, r, count=14):
    """Inputs of oracle/hostmpi/reduce_golden.c:fill for dtype t, rank r."""
    name = orc.DTYPES[t]
    i = np.arange(count)
    sval = ((r * 7 + i * 5) % 13) - 6
    uval = (r * 7 + i * 5) % 13
    if name == "MPI_C_BOOL":
        return (((r + i) % 3) == 0).astype(np.uint8)
    if name in ("MPI_BYTE", "MPI_CHAR"):
        return ((r * 37 + i * 11) & 0xFF).astype(np.uint8)
    npdt = np.dtype(orc.NP_DTYPE[name])
    if npdt.kind == "c":
        im = ((r * 3 + i * 2) % 7) - 3
        return (sval + 1j * im).astype(npdt)
    if npdt.kind == "f":
        return sval.astype(npdt)
    if npdt.kind == "u":
        return uval.astype(npdt)
    return sval.astype(npdt)


def _read_golden():
    recs = []
    with open(os.path.join(GOLDEN, "reduce_mpich.bin"), "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        t, op, count, nbytes = struct.unpack_from("<4i", data, off)
        off += 16
        recs.append((t, op, count, data[off:off + nbytes]))
        off += nbytes
    return recs


def test_reduce_fold_matches_mpich_for_every_legal_pair(orc):
    recs = _read_golden()
    legal = {(t, o) for t in range(30) for o in range(12) if orc.load().orc_op_allowed(t, o)}
    assert {(t, o) for t, o, _, _ in recs} == legal
    assert len(recs) == 216
    for t, op, count, raw in recs:
        ins = np.stack([_golden_inputs(orc, t, r, count) for r in range(5)])
        if orc.DTYPES[t] in ("MPI_C_BOOL", "MPI_BYTE"):
            ins = ins.astype(np.uint8)
        got = orc.reduce_fold(orc.DTYPES[t], orc.OPS[op], ins)
        assert got.tobytes() == raw, (orc.DTYPES[t], orc.OPS[op])


def test_legality_table_rejects(orc):
    for op in ("MPI_MAXLOC", "MPI_MINLOC"):
        for t in orc.DTYPES:
            assert not orc.op_allowed(t, op)
    assert not orc.op_allowed("MPI_CHAR", "MPI_SUM")
    assert not orc.op_allowed("MPI_FLOAT", "MPI_BAND")
    assert not orc.op_allowed("MPI_C_BOOL", "MPI_SUM")
    assert orc.op_allowed("MPI_BYTE", "MPI_BXOR")
    with pytest.raises(ValueError):
        orc.reduce_fold("MPI_FLOAT", "MPI_LXOR", np.zeros((2, 3), np.float32))


def test_reduce_closed_forms(orc):
    # Allreduce SUM of (r+1)(i+1) over 7 ranks (tests/collectives_test.cu:57-98).
    P, n = 7, 17
    ins = np.array([[(r + 1) * (i + 1) for i in range(n)] for r in range(P)], dtype=np.float64)
    got = orc.reduce_fold("MPI_DOUBLE", "MPI_SUM", ins)
    assert np.array_equal(got, np.array([(i + 1) * P * (P + 1) / 2 for i in range(n)]))
    # int16 PROD wraps around (tests/collectives_test.cu:6-54), 5 ranks, 14 elems.
    P, n = 5, 14
    ins = np.array([[(r + 1) * (i + 1) for i in range(n)] for r in range(P)], dtype=np.int16)
    got = orc.reduce_fold("MPI_INT16_T", "MPI_PROD", ins)
    exp = []
    for i in range(n):
        v = 1
        for r in range(P):
            v = (v * (r + 1) * (i + 1)) & 0xFFFF
        exp.append(v - 0x10000 if v >= 0x8000 else v)
    assert got.tolist() == exp


def test_fp_fold_is_rank_ordered(orc):
    # The fold is acc = x_r + acc in rank order: on inputs where order matters
    # it equals the sequential left fold and differs from the reversed fold.
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((64, 1000)) * 10.0 ** rng.integers(-8, 8, (64, 1))).astype(np.float32)
    got = orc.reduce_fold("MPI_FLOAT", "MPI_SUM", x)
    seq = x[0].copy()
    for r in range(1, 64):
        seq = (x[r] + seq).astype(np.float32)
    assert np.array_equal(got, seq)
    rev = x[-1].copy()
    for r in range(62, -1, -1):
        rev = (x[r] + rev).astype(np.float32)
    assert not np.array_equal(got, rev)


def test_max_min_operand_order(orc):
    # MAX(a=x_r, b=acc) = a > b ? a : b keeps acc on ties, so -0.0 vs +0.0
    # resolves to the earlier rank's value (operators.cu:73-81).
    x = np.array([[0.0], [-0.0]], dtype=np.float64)
    assert math.copysign(1, orc.reduce_fold("MPI_DOUBLE", "MPI_MAX", x)[0]) == 1.0
    x = np.array([[-0.0], [0.0]], dtype=np.float64)
    assert math.copysign(1, orc.reduce_fold("MPI_DOUBLE", "MPI_MAX", x)[0]) == -1.0


# --------------------------------------------------------------------------
# pi (toolchain_tests/pi/cpi.c)
# --------------------------------------------------------------------------
def test_pi_closed_form_and_host_mpi(orc):
    gold = json.load(open(os.path.join(GOLDEN, "cpi_mpich.json")))
    n = gold["n"]
    for p, printed in gold["pi_printed"].items():
        pi, parts = orc.pi_fold(n, int(p))
        assert abs(pi - math.pi) < 2e-12  # FP64 rounding of a 1e8-term sum
        # host MPI folds in another order: agreement to the FP64 tolerance
        assert abs(pi - float(printed)) / math.pi <= 1e-13
    # partial sums cover every interval exactly once: with one rank the fold
    # is the sequential midpoint sum itself
    pi1, parts1 = orc.pi_fold(1000, 1)
    h = 1.0 / 1000
    s = 0.0
    for i in range(1, 1001):
        x = h * (i - 0.5)
        s += 4.0 / (1.0 + x * x)
    assert pi1 == h * s


# --------------------------------------------------------------------------
# NPB IS: the benchmark's own known-answer test (is.c:254-282, 869-913,
# 1137): passed_verification == 5*MAX_ITERATIONS + comm_size.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("cls,nranks", [("S", 1), ("S", 2), ("S", 4), ("S", 16), ("W", 4), ("W", 32)])
def test_is_known_answers(orc, cls, nranks):
    res, keys, totals = orc.is_run(cls, nranks, want_keys=True)
    assert res.passed_verification == 50 + nranks
    assert res.full_verify_errors == 0
    assert res.total_keys_after == keys.size
    # Sorted output is a permutation of the generated keys (after the
    # iteration-dependent modifications of rank 0's first 20 keys).
    gen = np.concatenate([orc.is_keys(cls, nranks, r) for r in range(nranks)])
    mk = {"S": 1 << 11, "W": 1 << 16}[cls]
    for it in range(1, 11):
        gen[it] = it
        gen[it + 10] = mk - it
    assert np.array_equal(np.sort(gen), keys)


@pytest.mark.slow
def test_is_class_b_128(orc):
    res, _, totals = orc.is_run("B", 128)
    assert res.passed_verification == 50 + 128
    assert res.full_verify_errors == 0
    assert res.total_keys_after == 1 << 25


def test_randlc_is_the_textbook_lcg(orc):
    # x_{k+1} = 5^13 x_k mod 2^46, returned as 2^-46 x_{k+1}.
    import ctypes
    lib = orc.load()
    x = ctypes.c_double(314159265.0)
    a = 1220703125.0
    xi = 314159265
    for _ in range(1000):
        v = lib.orc_randlc(ctypes.byref(x), a)
        xi = (xi * 1220703125) % (1 << 46)
        assert x.value == float(xi)
        assert v == xi / float(1 << 46)


# --------------------------------------------------------------------------
# Payload generator: splitmix64 published vector (seed 1234567).
# --------------------------------------------------------------------------
def test_splitmix64_vector(orc):
    import ctypes
    lib = orc.load()
    s = ctypes.c_uint64(1234567)
    out = [lib.orc_splitmix64_next(ctypes.byref(s)) for _ in range(3)]
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_payload_bytes_are_little_endian_words(orc):
    lib = orc.load()
    b = orc.payload_bytes(7, 3, 20)
    w0, w1, w2 = (lib.orc_payload_word(7, 3, i) for i in range(3))
    exp = (w0.to_bytes(8, "little") + w1.to_bytes(8, "little") + w2.to_bytes(8, "little"))[:20]
    assert b.tobytes() == exp


# --------------------------------------------------------------------------
# Comm_split: tests/communicator_test.cu:93-152 known-answer table.
# --------------------------------------------------------------------------
def test_comm_split_known_answer(orc):
    colors, keys = [], []
    for r in range(10):
        if r < 3:
            colors.append(0), keys.append(r)
        elif r < 6:
            colors.append(1), keys.append(6 - r)
        elif r < 9:
            colors.append(2), keys.append(1)
        else:
            colors.append(2), keys.append(0)
    nr, ns = orc.comm_split(colors, keys)
    for r in range(10):
        if r < 3:
            assert (nr[r], ns[r]) == (r, 3)
        elif r < 6:
            assert (nr[r], ns[r]) == (5 - r, 3)
        elif r < 9:
            assert (nr[r], ns[r]) == (r - 5, 4)
        else:
            assert (nr[r], ns[r]) == (0, 4)


# --------------------------------------------------------------------------
# Wire codecs against torch's CPU casts.
# --------------------------------------------------------------------------
def test_bf16_matches_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 4000),
                        np.array([0.0, -0.0, np.inf, -np.inf, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8,
                                  3.3895314e38], dtype=np.float32)]).astype(np.float32)
    ours = orc.bf16_encode(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref)


def test_e4m3_matches_torch(orc):
    torch = pytest.importorskip("torch")
    if not hasattr(torch, "float8_e4m3fn"):
        pytest.skip("torch without float8_e4m3fn")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-448, 448, 3000), rng.standard_normal(3000) * 0.01,
                        np.array([0.0, 2 ** -9, 2 ** -10, 3 * 2 ** -10, 448.0, 0.015625, 1.0625, 1.1875])
                        ]).astype(np.float32)
    ours = orc.e4m3_encode(x)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(ours, ref)
    dec = orc.e4m3_decode(np.arange(256, dtype=np.uint8))
    tdec = torch.arange(256, dtype=torch.int32).to(torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    assert np.array_equal(np.isnan(dec), np.isnan(tdec))
    m = ~np.isnan(dec)
    assert np.array_equal(dec[m], tdec[m])
    # saturation beyond the format's range (torch itself maps these to NaN)
    assert orc.e4m3_encode(np.array([1e6, -500.0], dtype=np.float32)).tolist() == [0x7E, 0xFE]


def test_e4m3_block_roundtrip_error(orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(1 << 14).astype(np.float32)
    q, s, y = orc.e4m3_block_roundtrip(x, 256)
    rel = np.linalg.norm(y.astype(np.float64) - x) / np.linalg.norm(x.astype(np.float64))
    assert 1e-3 < rel < 0.05
    assert s.size == 64


# ---- the particle-in-cell workload (oracle.h orc_pic_run) -----------------
def _sorted_rows(q):
    v = np.ascontiguousarray(q).view(np.uint32).reshape(-1, 4)
    return v[np.lexsort(v.T[::-1])]


@pytest.mark.parametrize("shape", [(2, 2, 16, 16, 2000, 20), (3, 2, 8, 12, 500, 30), (1, 1, 16, 16, 300, 10),
                                   (2, 1, 12, 8, 400, 25), (4, 4, 32, 32, 5000, 10)])
def test_pic_decomposition_equals_the_global_grid(orc, shape):
    # The decomposed run (per-rank grids, stable partition, 8-direction
    # Sendrecv exchange, two-cell halos) must move every particle exactly as
    # the same physics on one undecomposed periodic grid: the same particle
    # states bit for bit (as a multiset), the same count, and energies that
    # differ only by FP64 summation order.
    px, py, nx, ny, np_, cycles = shape
    p = orc.pic_params(px, py, nx, ny, np_, cycles)
    d = orc.pic_run(p, want_particles=True)
    eg, qg = orc.pic_global(p)
    assert d["counts"].sum() == px * py * np_
    assert np.array_equal(_sorted_rows(d["particles"]), _sorted_rows(qg))
    assert np.max(np.abs(d["energy"] - eg) / eg) < 1e-13
    # BF16 wire: within the codec's relative error of the exact sum
    assert np.max(np.abs(d["energy_bf16"] - d["energy"]) / d["energy"]) < 1e-2


def test_pic_ballistic_closed_form(orc):
    # qm = 0: no force, so every velocity keeps its initial value (the
    # energy is that of the initial draws) and every position advances by
    # its velocity, wrapped.
    p = orc.pic_params(2, 2, 16, 16, 1000, 15, qm=0.0)
    d = orc.pic_run(p, want_particles=True)
    init = np.zeros((4 * 1000, 4), dtype=np.float32)
    buf = np.zeros(4, dtype=np.float32)
    for r in range(4):
        for i in range(1000):
            orc.load().orc_pic_init_particle(ctypes.byref(p), r, i, buf.ctypes.data_as(ctypes.c_void_p))
            init[r * 1000 + i] = buf
    e0 = sum(0.5 * (float(u) * float(u) + float(v) * float(v)) for u, v in init[:, 2:])
    assert np.allclose(d["energy"], e0, rtol=1e-13, atol=0)
    vel = lambda q: np.sort(q[:, 2:].view(np.uint64).ravel())
    assert np.array_equal(vel(d["particles"]), vel(init))


REF_IS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle", "_ref", "npb")
MPIRUN = "/opt/conda/bin/mpirun"


@pytest.mark.skipif(not (os.path.exists(os.path.join(REF_IS, "is.S.x")) and os.path.exists(MPIRUN)),
                    reason="the reference's NPB IS is built by `make -C oracle ref` where the reference is present")
@pytest.mark.parametrize("cls,P", [("S", 1), ("S", 4), ("S", 16), ("B", 8)])
def test_reference_npb_is_verifies_where_the_oracle_does(orc, cls, P):
    # The reference's own is.c under host MPI and the oracle's restatement
    # pass NPB's known-answer verification for the same class and ranks.
    import subprocess
    out = subprocess.run([MPIRUN, "-np", str(P), os.path.join(REF_IS, f"is.{cls}.x")], capture_output=True,
                         text=True, timeout=300).stdout
    assert "Verification    =               SUCCESSFUL" in out, out[-1500:]
    assert f"Size:  {33554432 if cls == 'B' else 65536}" in out.replace("  (class", " (class")
    res, _, _ = orc.is_run(cls, P)
    assert res.passed_verification == 5 * 10 + P
