"""NPB IS (integer bucket sort) on the device runtime against the CPU
oracle: the benchmark's own known-answer verification
(passed == 5 * MAX_ITERATIONS + comm_size, is.c:1137) and the globally
sorted keys bit-exact with the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOTAL = {"S": 1 << 16, "W": 1 << 20, "A": 1 << 23, "B": 1 << 25}


# Classes A and B get room in LDS for the ranking counts (the program's LDS
# path for the middle ranks, LDS windows for the edge ranks' wide key
# ranges); S and W keep only the bucket counters (its global-memory path).
USER_LDS = {"S": 4096, "W": 4096, "A": 4096 + 98304, "B": 4096 + 98304}
# The largest user LDS a 128-rank launch allows (the bench's setting).
USER_LDS_MAX_128 = 151552


def run_is(gm, cls, P, want_keys=True, threads=256, user_lds=None):
    out = gm.DeviceBuffer(8 * 8)
    out.zero()
    tout = gm.DeviceBuffer(8 * 4)
    keys = gm.DeviceBuffer(4 * TOTAL[cls]) if want_keys else None
    res = gm.run("npb_is", P, [0, ord(cls), out, tout, keys, 10], threads_per_rank=threads,
                 user_lds_bytes=user_lds or USER_LDS[cls], heap_bytes=P * (32 << 20) if cls == "B" else P * (64 << 20) // max(1, P // 8),
                 max_requests=64, timeout_s=120.0)
    assert res.status == 0, res.describe()
    o = out.download(np.int64)
    t = tout.download(np.float64)
    k = keys.download(np.int32) if keys is not None else None
    return o, t, k


@pytest.mark.parametrize("cls,P", [("S", 1), ("S", 2), ("S", 16), ("W", 4), ("W", 32), ("A", 8)])
def test_is_matches_oracle(gm, orc, cls, P):
    print(f"[gpu] npb_is class {cls} P={P}", flush=True)
    o, t, keys = run_is(gm, cls, P)
    assert o[0] == 50 + P, o
    assert o[2] == 0
    assert o[1] == TOTAL[cls]
    res, okeys, _ = orc.is_run(cls, P, want_keys=True)
    assert res.passed_verification == 50 + P
    assert np.array_equal(keys, okeys)


@pytest.mark.parametrize("user_lds", [USER_LDS["B"], USER_LDS_MAX_128])
def test_is_class_b_128(gm, orc, user_lds):
    print(f"[gpu] npb_is class B P=128 user_lds={user_lds}", flush=True)
    o, t, keys = run_is(gm, "B", 128, user_lds=user_lds)
    assert o[0] == 50 + 128, o
    assert o[2] == 0 and o[1] == 1 << 25
    res, okeys, _ = orc.is_run("B", 128, want_keys=True)
    assert np.array_equal(keys, okeys)
    print(f"[gpu] class B: {10 * (1 << 25) / t[0] / 1e6:.1f} Mkeys/s, alltoallv {t[1]:.4f}s", flush=True)
