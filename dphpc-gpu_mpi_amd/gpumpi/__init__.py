"""Python host binding of the gpumpi C-ABI (include/gpumpi/host.h).

Mirrors the reference's host entry points for this path: the `---gpumpi`
launcher of gpu_libs/gpu_main/main.cu:133-321 (here `run()` /
`Launcher.run_argv`) and the test launcher tests/test_runner.cuh:26-86
(here `run()` with a per-rank `ok` array).  Programs are rank programs
compiled into lib/libgpumpi_progs.so and looked up by name.

Importing this module never falls back to anything: if the shared
libraries are missing, `load()` raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
CORE_LIB = os.path.join(LIB_DIR, "libgpumpi.so")
PROGS_LIB = os.path.join(LIB_DIR, "libgpumpi_progs.so")

# Datatype enum (gpu_libs/gpu_mpi/mpi_types_list.cuh:3-33).
DTYPES = ["MPI_CHAR", "MPI_SHORT", "MPI_INT", "MPI_LONG", "MPI_LONG_LONG_INT",
          "MPI_LONG_LONG", "MPI_SIGNED_CHAR", "MPI_UNSIGNED_CHAR", "MPI_UNSIGNED_SHORT",
          "MPI_UNSIGNED", "MPI_UNSIGNED_LONG", "MPI_UNSIGNED_LONG_LONG", "MPI_FLOAT",
          "MPI_DOUBLE", "MPI_LONG_DOUBLE", "MPI_WCHAR", "MPI_C_BOOL", "MPI_INT8_T",
          "MPI_INT16_T", "MPI_INT32_T", "MPI_INT64_T", "MPI_UINT8_T", "MPI_UINT16_T",
          "MPI_UINT32_T", "MPI_UINT64_T", "MPI_C_COMPLEX", "MPI_C_FLOAT_COMPLEX",
          "MPI_C_DOUBLE_COMPLEX", "MPI_C_LONG_DOUBLE_COMPLEX", "MPI_BYTE"]
DT = {n: i for i, n in enumerate(DTYPES)}
OPS = ["MPI_MAX", "MPI_MIN", "MPI_SUM", "MPI_PROD", "MPI_LAND", "MPI_BAND", "MPI_LOR",
       "MPI_BOR", "MPI_LXOR", "MPI_BXOR", "MPI_MAXLOC", "MPI_MINLOC"]
OP = {n: i for i, n in enumerate(OPS)}

# Exported symbols of each library (the C-ABI surface).
CORE_SYMBOLS = ["gpumpi_default_config", "gpumpi_parse_args", "gpumpi_run", "gpumpi_run_graph", "gpumpi_run_noalloc",
                "gpumpi_main",
                "gpumpi_device_alloc", "gpumpi_device_free", "gpumpi_memcpy_htod",
                "gpumpi_memcpy_dtoh", "gpumpi_memset", "gpumpi_max_ranks", "gpumpi_last_error",
                "gpumpi_reset_world", "gpumpi_release_world", "gpumpi_service_poll", "gpumpi_apply_tuning"]
PROGS_SYMBOLS = ["gpumpi_prog_lookup", "gpumpi_prog_count", "gpumpi_prog_name",
                 "gpumpi_fill_payload", "gpumpi_verify_payload", "gpumpi_fill_normal",
                 "gpumpi_copy_calibration"]

STATUS = {0: "OK", 1: "EINVAL", 2: "EHIP", 3: "ETOOBIG", 4: "EDEVICE", 5: "ETIMEOUT", 6: "ENOMEM"}


class Config(ctypes.Structure):
    _fields_ = [("num_procs", ctypes.c_int), ("threads_per_rank", ctypes.c_int),
                ("blocks_per_grid", ctypes.c_int), ("max_requests", ctypes.c_int),
                ("chan_slots", ctypes.c_int), ("eager_bytes", ctypes.c_int),
                ("helper_min_bytes", ctypes.c_int), ("xcd_map", ctypes.c_int),
                ("stack_size", ctypes.c_size_t), ("heap_bytes", ctypes.c_size_t),
                ("staging_bytes", ctypes.c_size_t), ("comm_pool_bytes", ctypes.c_size_t),
                ("timeout_s", ctypes.c_double), ("verbose", ctypes.c_int),
                ("user_lds_bytes", ctypes.c_int), ("reserved", ctypes.c_int * 6)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("any_rank_failed", ctypes.c_int),
                ("mpi_error", ctypes.c_int), ("mpi_error_rank", ctypes.c_int),
                ("mpi_error_line", ctypes.c_int), ("mpi_error_count", ctypes.c_int),
                ("timed_out", ctypes.c_int), ("lds_bytes", ctypes.c_int),
                ("kernel_ms", ctypes.c_double), ("mpi_error_msg", ctypes.c_char * 112)]

    def describe(self):
        s = f"status={STATUS.get(self.status, self.status)} kernel_ms={self.kernel_ms:.3f}"
        if self.mpi_error:
            s += (f" mpi_error={self.mpi_error} rank={self.mpi_error_rank} line={self.mpi_error_line}"
                  f" msg={self.mpi_error_msg.decode(errors='replace')!r}")
        return s


class GpumpiError(RuntimeError):
    pass


_core = None
_progs = None


def libs_built():
    return os.path.exists(CORE_LIB) and os.path.exists(PROGS_LIB)


def load_core():
    """The host C-ABI library alone (no GPU needed to load it)."""
    global _core
    if _core is None:
        if not os.path.exists(CORE_LIB):
            raise GpumpiError(f"{CORE_LIB} is missing: build it with dphpc-gpu_mpi_amd/build.py")
        lib = ctypes.CDLL(CORE_LIB, mode=ctypes.RTLD_GLOBAL)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        lib.gpumpi_default_config.argtypes = [ctypes.POINTER(Config), ctypes.c_int]
        lib.gpumpi_parse_args.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(Config)]
        lib.gpumpi_apply_tuning.argtypes = [ctypes.POINTER(Config), ctypes.c_char_p]
        lib.gpumpi_run.argtypes = [ctypes.POINTER(Config), vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                   vp, ctypes.POINTER(Result)]
        lib.gpumpi_run_graph.argtypes = [ctypes.POINTER(Config), vp, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                         vp, ctypes.c_int, ctypes.POINTER(Result), vp]
        lib.gpumpi_run_noalloc.argtypes = [ctypes.POINTER(Config), vp, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_char_p), vp, ctypes.POINTER(Result)]
        lib.gpumpi_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), vp]
        lib.gpumpi_device_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), sz]
        lib.gpumpi_device_free.argtypes = [vp]
        lib.gpumpi_memcpy_htod.argtypes = [vp, vp, sz]
        lib.gpumpi_memcpy_dtoh.argtypes = [vp, vp, sz]
        lib.gpumpi_memset.argtypes = [vp, ctypes.c_int, sz]
        lib.gpumpi_max_ranks.argtypes = [ctypes.POINTER(Config), vp]
        lib.gpumpi_last_error.restype = ctypes.c_char_p
        _core = lib
    return _core


def load():
    """Both libraries; raises if they are not built."""
    global _progs
    core = load_core()
    if _progs is None:
        if not os.path.exists(PROGS_LIB):
            raise GpumpiError(f"{PROGS_LIB} is missing: build it with dphpc-gpu_mpi_amd/build.py")
        lib = ctypes.CDLL(PROGS_LIB, mode=ctypes.RTLD_GLOBAL)
        lib.gpumpi_prog_lookup.argtypes = [ctypes.c_char_p]
        lib.gpumpi_prog_lookup.restype = ctypes.c_void_p
        lib.gpumpi_prog_name.argtypes = [ctypes.c_int]
        lib.gpumpi_prog_name.restype = ctypes.c_char_p
        lib.gpumpi_fill_payload.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_int, ctypes.c_int]
        lib.gpumpi_verify_payload.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.gpumpi_verify_payload.restype = ctypes.c_longlong
        lib.gpumpi_fill_normal.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.gpumpi_copy_calibration.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        _progs = lib
    return core, _progs


def reset_world():
    load_core().gpumpi_reset_world()


def release_world():
    """Free the device world kept between launches."""
    load_core().gpumpi_release_world()


def last_error():
    return load_core().gpumpi_last_error().decode(errors="replace")


def programs():
    _, p = load()
    return [p.gpumpi_prog_name(i).decode() for i in range(p.gpumpi_prog_count())]


def default_config(nranks, tuning=None, **kw):
    """gpumpi_default_config, then a tuning file's choices (gpumpi_apply_tuning,
    written by gpumpi.tune), then the keyword overrides."""
    core = load_core()
    c = Config()
    core.gpumpi_default_config(ctypes.byref(c), nranks)
    if tuning is not None and core.gpumpi_apply_tuning(ctypes.byref(c), os.fsencode(tuning)) != 0:
        raise GpumpiError(f"tuning file {tuning}: {core.gpumpi_last_error().decode()}")
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


# Live device buffers.  At interpreter exit they are freed, and the world
# released, from an atexit handler -- while the HIP runtime is still up --
# rather than from __del__ during module teardown, which can run after the
# runtime's own teardown has begun.
_live = weakref.WeakSet()
_shutting_down = False


def _release_all():
    global _shutting_down
    if _core is None:
        return
    for b in list(_live):
        b.free()
    _core.gpumpi_release_world()
    _shutting_down = True


atexit.register(_release_all)


class DeviceBuffer:
    """A device allocation made through the C-ABI."""

    def __init__(self, nbytes):
        core = load_core()
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        if core.gpumpi_device_alloc(ctypes.byref(p), max(self.nbytes, 1)) != 0:
            raise GpumpiError(last_error())
        self.ptr = p.value
        _live.add(self)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        if load_core().gpumpi_memcpy_htod(self.ptr + offset, a.ctypes.data, a.nbytes) != 0:
            raise GpumpiError(last_error())

    def download(self, dtype=np.uint8, count=None, offset=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if load_core().gpumpi_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes) != 0:
            raise GpumpiError(last_error())
        return out

    def zero(self, value=0):
        if load_core().gpumpi_memset(self.ptr, value, self.nbytes) != 0:
            raise GpumpiError(last_error())

    def free(self):
        if self.ptr and not _shutting_down:
            load_core().gpumpi_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def run(program, nranks, args=(), check=True, argv=(), graph_replays=0, noalloc=False, **cfg_kw):
    """Launch rank program `program` on `nranks` ranks.

    `args` are 64-bit words handed to the program (ints or DeviceBuffers,
    which pass their device address).  With graph_replays > 0 the launch is
    captured in a hipGraph and replayed that many times (Result.kernel_ms is
    the mean replay time; Result.replay_ms the list).  With noalloc the
    launch goes through gpumpi_run_noalloc (the world in the library's static
    device array).  Returns the Result.
    """
    core, progs = load()
    kern = progs.gpumpi_prog_lookup(program.encode())
    if not kern:
        raise GpumpiError(f"no rank program named {program!r}")
    cfg = default_config(nranks, **cfg_kw)
    words = np.zeros(max(64, len(args)), dtype=np.uint64)
    for i, a in enumerate(args):
        if isinstance(a, DeviceBuffer):
            words[i] = a.ptr
        elif a is None:
            words[i] = 0
        else:
            words[i] = np.uint64(int(a) & 0xFFFFFFFFFFFFFFFF)
    dargs = DeviceBuffer.from_array(words)
    cargv = (ctypes.c_char_p * (len(argv) + 1))(*[s.encode() for s in argv], None)
    res = Result()
    trace = os.environ.get("GPUMPI_TRACE")
    if trace:
        print(f"[gpumpi] launch {program} P={nranks} threads={cfg.threads_per_rank} "
              f"replays={graph_replays} args={[int(w) for w in words[:len(args)]]}", flush=True)
    if graph_replays > 0:
        ms = np.zeros(graph_replays, dtype=np.float64)
        rc = core.gpumpi_run_graph(ctypes.byref(cfg), kern, len(argv), cargv, dargs.ptr, graph_replays,
                                   ctypes.byref(res), ms.ctypes.data)
        res.replay_ms = ms.tolist()
    elif noalloc:
        rc = core.gpumpi_run_noalloc(ctypes.byref(cfg), kern, len(argv), cargv, dargs.ptr, ctypes.byref(res))
    else:
        rc = core.gpumpi_run(ctypes.byref(cfg), kern, len(argv), cargv, dargs.ptr, ctypes.byref(res))
    err = last_error() if rc != 0 else ""  # before any later call can overwrite it
    dargs.free()
    if trace:
        print(f"[gpumpi] done {program} rc={rc} status={res.status} {err}", flush=True)
    if check and rc != 0:
        raise GpumpiError(f"{program}: {res.describe()} :: {err}")
    return res


def fill_payload(buf, stride, nbytes, seed, first_rank, nbuf):
    _, progs = load()
    if progs.gpumpi_fill_payload(buf.ptr, stride, nbytes, seed, first_rank, nbuf) != 0:
        raise GpumpiError("fill_payload failed")


def fill_normal(buf, n, seed, stream, offset_elems=0):
    _, progs = load()
    if progs.gpumpi_fill_normal(buf.ptr + 4 * offset_elems, n, seed, stream) != 0:
        raise GpumpiError("fill_normal failed")


def copy_calibration(dst, src, nbytes):
    """Plain device copy of nbytes (multiple of 16): the reference kernel for
    HBM traffic counters."""
    _, progs = load()
    if progs.gpumpi_copy_calibration(dst.ptr, src.ptr, nbytes) != 0:
        raise GpumpiError("copy_calibration failed")


def verify_payload(buf, stride, nbytes, seed, shift, nranks, nbuf):
    _, progs = load()
    return int(progs.gpumpi_verify_payload(buf.ptr, stride, nbytes, seed, shift, nranks, nbuf))
