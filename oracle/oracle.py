"""ctypes front end of the CPU oracle (oracle.c) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, the smoke check and bench.py's CPU-baseline leg, never by
the product path.  `load()` builds liboracle.so with oracle/Makefile when it
is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

# Datatype enum, gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33.
DTYPES = ["MPI_CHAR", "MPI_SHORT", "MPI_INT", "MPI_LONG", "MPI_LONG_LONG_INT",
          "MPI_LONG_LONG", "MPI_SIGNED_CHAR", "MPI_UNSIGNED_CHAR", "MPI_UNSIGNED_SHORT",
          "MPI_UNSIGNED", "MPI_UNSIGNED_LONG", "MPI_UNSIGNED_LONG_LONG", "MPI_FLOAT",
          "MPI_DOUBLE", "MPI_LONG_DOUBLE", "MPI_WCHAR", "MPI_C_BOOL", "MPI_INT8_T",
          "MPI_INT16_T", "MPI_INT32_T", "MPI_INT64_T", "MPI_UINT8_T", "MPI_UINT16_T",
          "MPI_UINT32_T", "MPI_UINT64_T", "MPI_C_COMPLEX", "MPI_C_FLOAT_COMPLEX",
          "MPI_C_DOUBLE_COMPLEX", "MPI_C_LONG_DOUBLE_COMPLEX", "MPI_BYTE"]
DT = {n: i for i, n in enumerate(DTYPES)}
# Operator enum, gpu_libs/gpu_mpi/mpi_operators_list.cuh:4-16.
OPS = ["MPI_MAX", "MPI_MIN", "MPI_SUM", "MPI_PROD", "MPI_LAND", "MPI_BAND", "MPI_LOR",
       "MPI_BOR", "MPI_LXOR", "MPI_BXOR", "MPI_MAXLOC", "MPI_MINLOC"]
OP = {n: i for i, n in enumerate(OPS)}

# numpy dtype of each MPI datatype (complex as interleaved pairs).
NP_DTYPE = {
    "MPI_CHAR": np.int8, "MPI_SHORT": np.int16, "MPI_INT": np.int32, "MPI_LONG": np.int64,
    "MPI_LONG_LONG_INT": np.int64, "MPI_LONG_LONG": np.int64, "MPI_SIGNED_CHAR": np.int8,
    "MPI_UNSIGNED_CHAR": np.uint8, "MPI_UNSIGNED_SHORT": np.uint16, "MPI_UNSIGNED": np.uint32,
    "MPI_UNSIGNED_LONG": np.uint64, "MPI_UNSIGNED_LONG_LONG": np.uint64,
    "MPI_FLOAT": np.float32, "MPI_DOUBLE": np.float64, "MPI_LONG_DOUBLE": np.float64,
    "MPI_WCHAR": np.int32, "MPI_C_BOOL": np.uint8, "MPI_INT8_T": np.int8,
    "MPI_INT16_T": np.int16, "MPI_INT32_T": np.int32, "MPI_INT64_T": np.int64,
    "MPI_UINT8_T": np.uint8, "MPI_UINT16_T": np.uint16, "MPI_UINT32_T": np.uint32,
    "MPI_UINT64_T": np.uint64, "MPI_C_COMPLEX": np.complex64,
    "MPI_C_FLOAT_COMPLEX": np.complex64, "MPI_C_DOUBLE_COMPLEX": np.complex128,
    "MPI_C_LONG_DOUBLE_COMPLEX": np.complex128, "MPI_BYTE": np.uint8,
}

_lib = None


class ISResult(ctypes.Structure):
    _fields_ = [("passed_verification", ctypes.c_int),
                ("full_verify_errors", ctypes.c_int),
                ("total_keys_after", ctypes.c_int64),
                ("sorted_checksum", ctypes.c_uint64),
                ("timed_seconds", ctypes.c_double)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    src = os.path.join(HERE, "oracle.c")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = ctypes.CDLL(LIB)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.orc_type_size.argtypes = [ctypes.c_int]
    lib.orc_op_allowed.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.orc_reduce_fold.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, sz, vp, vp]
    lib.orc_pi_partial.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_pi_partial.restype = ctypes.c_double
    lib.orc_pi_fold.argtypes = [ctypes.c_int, ctypes.c_int, vp]
    lib.orc_pi_fold.restype = ctypes.c_double
    lib.orc_mix64.argtypes = [ctypes.c_uint64]
    lib.orc_mix64.restype = ctypes.c_uint64
    lib.orc_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
    lib.orc_splitmix64_next.restype = ctypes.c_uint64
    lib.orc_payload_word.argtypes = [ctypes.c_uint64] * 3
    lib.orc_payload_word.restype = ctypes.c_uint64
    lib.orc_payload_bytes.argtypes = [ctypes.c_uint64, ctypes.c_uint64, sz, vp]
    lib.orc_randlc.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_double]
    lib.orc_randlc.restype = ctypes.c_double
    lib.orc_find_my_seed.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long,
                                     ctypes.c_double, ctypes.c_double]
    lib.orc_find_my_seed.restype = ctypes.c_double
    lib.orc_is_run.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.POINTER(ISResult), vp, vp]
    lib.orc_is_keys.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int, vp]
    lib.orc_f32_to_bf16.argtypes = [ctypes.c_float]
    lib.orc_f32_to_bf16.restype = ctypes.c_uint16
    lib.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
    lib.orc_bf16_to_f32.restype = ctypes.c_float
    lib.orc_f32_to_e4m3.argtypes = [ctypes.c_float]
    lib.orc_f32_to_e4m3.restype = ctypes.c_uint8
    lib.orc_e4m3_to_f32.argtypes = [ctypes.c_uint8]
    lib.orc_e4m3_to_f32.restype = ctypes.c_float
    lib.orc_e4m3_block_encode.argtypes = [vp, sz, ctypes.c_int, vp, vp]
    lib.orc_e4m3_block_decode.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    lib.orc_comm_split.argtypes = [ctypes.c_int, vp, vp, vp, vp]
    lib.orc_pic_run.argtypes = [ctypes.c_void_p, vp, vp, vp, vp, vp, vp]
    lib.orc_pic_global.argtypes = [ctypes.c_void_p, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    lib.orc_pic_energy.argtypes = [vp, ctypes.c_int64]
    lib.orc_pic_energy.restype = ctypes.c_double
    lib.orc_pic_init_particle.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, vp]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def reduce_fold(dtype, op, inputs):
    """inputs: array of shape (P, count) of NP_DTYPE[dtype]; rank-order fold."""
    lib = load()
    inputs = np.ascontiguousarray(inputs)
    out = np.empty(inputs.shape[1:], dtype=inputs.dtype)
    rc = lib.orc_reduce_fold(DT[dtype], OP[op], inputs.shape[0], inputs.shape[1], _p(inputs), _p(out))
    if rc != 0:
        raise ValueError(f"illegal reduction {dtype} {op}")
    return out


def op_allowed(dtype, op):
    return bool(load().orc_op_allowed(DT[dtype], OP[op]))


def type_size(dtype):
    return load().orc_type_size(DT[dtype])


def pi_fold(n, nranks):
    parts = np.zeros(nranks, dtype=np.float64)
    pi = load().orc_pi_fold(n, nranks, _p(parts))
    return pi, parts


def payload_bytes(seed, rank, nbytes):
    out = np.empty(nbytes, dtype=np.uint8)
    load().orc_payload_bytes(seed, rank, nbytes, _p(out))
    return out


def is_run(cls, nranks, want_keys=False):
    lib = load()
    res = ISResult()
    total = {"S": 1 << 16, "W": 1 << 20, "A": 1 << 23, "B": 1 << 25}[cls]
    nb = 512 if cls == "S" else 1024
    keys = np.zeros(total, dtype=np.int32) if want_keys else None
    totals = np.zeros(nb + 5, dtype=np.int32)
    rc = lib.orc_is_run(cls.encode(), nranks, ctypes.byref(res),
                        _p(keys) if keys is not None else None, _p(totals))
    if rc != 0:
        raise RuntimeError(f"orc_is_run failed rc={rc}")
    return res, keys, totals


def is_keys(cls, nranks, rank):
    total = {"S": 1 << 16, "W": 1 << 20, "A": 1 << 23, "B": 1 << 25}[cls]
    out = np.empty(total // nranks, dtype=np.int32)
    load().orc_is_keys(cls.encode(), nranks, rank, _p(out))
    return out


def bf16_encode(x):
    lib = load()
    x = np.asarray(x, dtype=np.float32).ravel()
    return np.array([lib.orc_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)


def e4m3_encode(x):
    lib = load()
    x = np.asarray(x, dtype=np.float32).ravel()
    return np.array([lib.orc_f32_to_e4m3(float(v)) for v in x], dtype=np.uint8)


def e4m3_decode(q):
    lib = load()
    return np.array([lib.orc_e4m3_to_f32(int(v)) for v in np.asarray(q).ravel()], dtype=np.float32)


def e4m3_block_roundtrip(x, block=256):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.size
    q = np.empty(n, dtype=np.uint8)
    s = np.empty((n + block - 1) // block, dtype=np.float32)
    lib.orc_e4m3_block_encode(_p(x), n, block, _p(q), _p(s))
    y = np.empty(n, dtype=np.float32)
    lib.orc_e4m3_block_decode(_p(q), _p(s), n, block, _p(y))
    return q, s, y


def comm_split(colors, keys):
    colors = np.ascontiguousarray(colors, dtype=np.int32)
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    n = colors.size
    nr = np.empty(n, dtype=np.int32)
    ns = np.empty(n, dtype=np.int32)
    load().orc_comm_split(n, _p(colors), _p(keys), _p(nr), _p(ns))
    return nr, ns


class PicParams(ctypes.Structure):
    _fields_ = [("px", ctypes.c_int), ("py", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("np", ctypes.c_int64), ("cycles", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("vth", ctypes.c_float), ("drift", ctypes.c_float), ("qm", ctypes.c_float)]


def pic_params(px, py, nx, ny, np_, cycles, seed=2024, vth=0.02, drift=0.01, qm=0.001):
    return PicParams(px, py, nx, ny, np_, cycles, seed, vth, drift, qm)


def pic_run(p, want_particles=False):
    """The decomposed PIC run: dict of energy, energy_bf16, counts,
    part_hash, rho_hash (per rank) and, optionally, the final particles."""
    lib = load()
    P = p.px * p.py
    e = np.zeros(p.cycles, dtype=np.float64)
    e16 = np.zeros(p.cycles, dtype=np.float32)
    counts = np.zeros(P, dtype=np.int64)
    ph = np.zeros(P, dtype=np.uint64)
    rh = np.zeros(P, dtype=np.uint64)
    q = np.zeros((P * p.np, 4), dtype=np.float32) if want_particles else None
    rc = lib.orc_pic_run(ctypes.byref(p), _p(e), _p(e16), _p(counts), _p(ph), _p(rh),
                         _p(q) if q is not None else None)
    if rc != 0:
        raise RuntimeError(f"orc_pic_run failed ({rc})")
    return {"energy": e, "energy_bf16": e16, "counts": counts, "part_hash": ph, "rho_hash": rh, "particles": q}


def pic_global(p):
    lib = load()
    n = p.px * p.py * p.np
    e = np.zeros(p.cycles, dtype=np.float64)
    q = np.zeros((n, 4), dtype=np.float32)
    nn = ctypes.c_int64(0)
    if lib.orc_pic_global(ctypes.byref(p), _p(e), _p(q), ctypes.byref(nn)) != 0:
        raise RuntimeError("orc_pic_global failed")
    return e, q
