"""The autotuner on the device (gpumpi/tune.py): a coordinate-descent pass
over threads per rank, copy-split threshold, eager threshold and channel
slots with the ring objective at 16 ranks; the chosen knobs written to a
tuning file, applied through gpumpi_apply_tuning, and a ring under them
delivering every payload intact."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "dphpc-gpu_mpi_amd"))


def test_autotune_pass_and_tuned_ring(gm, tmp_path):
    from gpumpi import tune
    P = 16
    print(f"[gpu] autotune P={P}", flush=True)
    res = tune.tune_device(gm, P, repeats=1, sizes=(8, 4096, 262144, 4 << 20),
                           log=lambda m: print(f"[tune] {m}", flush=True))
    best = res["best"]
    for key, cands in tune.KNOBS:
        assert best[key] in cands or best[key] == res["trials"][0]["knobs"][key]
    assert 0 < res["objective"] <= res["default_objective"]
    path = str(tmp_path / "gpu.tuning")
    tune.write_tuning(path, best)
    S = 1 << 20
    sb = gm.DeviceBuffer(S * P)
    rb = gm.DeviceBuffer(S * P)
    rb.zero()
    gm.fill_payload(sb, S, S, 77, 0, P)
    szb = gm.DeviceBuffer.from_array(np.array([S], dtype=np.uint64))
    itb = gm.DeviceBuffer.from_array(np.array([5], dtype=np.int64))
    wmb = gm.DeviceBuffer.from_array(np.array([1], dtype=np.int64))
    tb = gm.DeviceBuffer(8 * P)
    gm.run("p2p_ring_sweep", P, [0, szb, 1, itb, wmb, sb, rb, S, tb], tuning=path, timeout_s=60.0)
    assert gm.verify_payload(rb, S, S, 77, -1, P, P) == 0
