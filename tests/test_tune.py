"""The autotuner's search and its tuning file, on the CPU: coordinate
descent over a synthetic objective with a known optimum and refused
candidates, and the file applied by the C-ABI (gpumpi_apply_tuning) and the
`---gpumpi --tuning FILE` option."""
import ctypes
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "dphpc-gpu_mpi_amd"))

from gpumpi import tune  # noqa: E402

DEFAULTS = {"threads_per_rank": 256, "helper_min_bytes": 16384, "eager_bytes": 448, "chan_slots": 8}


def synthetic(best, refused=()):
    # separable, convex in log2 of every knob, minimum at `best`
    import math
    calls = []

    def measure(kn):
        calls.append(dict(kn))
        for k, v in refused:
            if kn[k] == v:
                raise RuntimeError(f"{k}={v} refused")
        return 1.0 + sum((math.log2(kn[k]) - math.log2(best[k])) ** 2 for k in best)
    return measure, calls


def test_coordinate_descent_finds_the_separable_optimum():
    best = {"threads_per_rank": 512, "helper_min_bytes": 65536, "eager_bytes": 1984, "chan_slots": 4}
    measure, calls = synthetic(best)
    res = tune.autotune(measure, DEFAULTS, repeats=1)
    assert res["best"] == best and res["objective"] == 1.0
    assert res["default_objective"] > 1.0
    # every candidate of every knob is tried once, plus the defaults
    assert len(calls) == 1 + sum(len(c) - 1 for _, c in tune.KNOBS)


def test_refused_candidates_are_skipped_and_recorded():
    best = {"threads_per_rank": 1024, "helper_min_bytes": 16384, "eager_bytes": 448, "chan_slots": 8}
    measure, _ = synthetic(best, refused=[("threads_per_rank", 1024)])
    res = tune.autotune(measure, DEFAULTS, repeats=2)
    assert res["best"]["threads_per_rank"] == 512  # the best one that launches
    skipped = [t for t in res["trials"] if t["objective"] is None]
    assert len(skipped) == 1 and "refused" in skipped[0]["error"]


def test_defaults_kept_when_nothing_is_better():
    measure, _ = synthetic(DEFAULTS)
    res = tune.autotune(measure, DEFAULTS, repeats=1)
    assert res["best"] == DEFAULTS and res["objective"] == res["default_objective"]


def test_tuning_file_round_trip_and_c_abi(tmp_path):
    import gpumpi as gm
    path = str(tmp_path / "t.tuning")
    vals = {"threads_per_rank": 512, "helper_min_bytes": 65536, "eager_bytes": 960, "chan_slots": 16}
    tune.write_tuning(path, vals, {"ranks": 256, "objective_s": 1e-5})
    assert tune.read_tuning(path) == vals
    c = gm.default_config(64, tuning=path)
    assert {k: int(getattr(c, k)) for k in vals} == vals
    # the CLI option applies the same file
    core = gm.load_core()
    cfg = gm.default_config(64)
    argv = [b"prog", b"---gpumpi", b"-n", b"8", b"--tuning", path.encode()]
    arr = (ctypes.c_char_p * len(argv))(*argv)
    assert core.gpumpi_parse_args(len(argv), arr, ctypes.byref(cfg)) == 1
    assert cfg.num_procs == 8 and cfg.eager_bytes == 960 and cfg.chan_slots == 16


@pytest.mark.parametrize("text", ["chan_slots 6\n", "threads_per_rank 100\n", "eager_bytes -1\n",
                                  "bogus_key 3\n", "eager_bytes\n", "helper_min_bytes 1 2\n"])
def test_bad_tuning_files_are_refused_and_leave_the_config(tmp_path, text):
    import gpumpi as gm
    path = str(tmp_path / "bad.tuning")
    with open(path, "w") as f:
        f.write("eager_bytes 960\n" + text)
    core = gm.load_core()
    cfg = gm.default_config(16)
    before = bytes(cfg)
    assert core.gpumpi_apply_tuning(ctypes.byref(cfg), path.encode()) != 0
    assert bytes(cfg) == before  # unchanged: nothing applied
    assert b"bad.tuning" in core.gpumpi_last_error()
    with pytest.raises(gm.GpumpiError):
        gm.default_config(16, tuning=path)
