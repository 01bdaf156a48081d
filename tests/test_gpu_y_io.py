"""MPI_File_* and the printf family on the GPU, through the host service
channel the launcher serves while the launch runs (the same scenarios and
expectations as tests/test_simt.py::test_sim_io_scenarios; see
tests/io_expect.py for what each record checks)."""
import os

import numpy as np
import pytest

import io_expect

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("P,N", [(1, 5), (4, 10), (64, 300)])
def test_io_scenarios(gm, tmp_path, capfd, P, N):
    print(f"[gpu] io_scenarios P={P} N={N}", flush=True)
    rec = gm.DeviceBuffer(4 * P * io_expect.KREC)
    rec.zero()
    txt = gm.DeviceBuffer(256 * P)
    txt.zero()
    path = np.frombuffer(str(tmp_path).encode() + b"\0", dtype=np.uint8)
    dirb = gm.DeviceBuffer.from_array(path)
    res = gm.run("io_scenarios", P, [rec, dirb, txt, N], threads_per_rank=128, timeout_s=60.0)
    assert res.status == 0, res.describe()
    recs = rec.download(np.int32).reshape(P, io_expect.KREC)
    raw = txt.download(np.uint8).reshape(P, 256)
    texts = [bytes(raw[r]).split(b"\0")[0].decode() for r in range(P)]
    io_expect.check(recs, texts, P, N)
    out = capfd.readouterr().out
    for r in range(P):
        assert f"io_scenarios: rank {r} of {P}" in out
    assert os.listdir(tmp_path) == []
