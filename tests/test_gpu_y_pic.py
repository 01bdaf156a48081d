"""The C5 particle-in-cell cycle (progs/pic_exact.hip) on the device runtime
against the oracle's restatement (orc_pic_run), bit for bit: every cycle's
FP64 energy (MPI_Allreduce, rank-ordered fold), and per rank the final
particle count, the FNV-1a hash of its particle array (x, y, u, v FP32) and
of its density with the two-cell halo ring.

`pic_exact` is the cycle without the BF16-wire energy Allreduce; the program
with it (`pic`) is covered by the CPU simulation (tests/test_simt.py)."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_pic(gm, px, py, nx, ny, np_, cycles, seed=2024, threads=256, want_phases=False):
    P = px * py
    cap = 2 * np_ + 1024
    e = gm.DeviceBuffer(8 * cycles)
    e16 = gm.DeviceBuffer(4 * cycles)
    cnt = gm.DeviceBuffer(8 * P)
    ph = gm.DeviceBuffer(8 * P)
    rh = gm.DeviceBuffer(8 * P)
    tp = gm.DeviceBuffer(8 * 7)
    for b in (e, e16, cnt, ph, rh, tp):
        b.zero()
    f = lambda x: struct.unpack("<I", struct.pack("<f", x))[0]  # noqa: E731
    heap = 16 * 3 * cap + cap + (2 << 20)
    res = gm.run("pic_exact", P, [0, px, py, nx, ny, np_, cycles, seed, f(0.02), f(0.01), f(0.001), e, e16, cnt, ph,
                                  rh, tp if want_phases else 0, cap],
                 threads_per_rank=threads, heap_bytes=P * heap, timeout_s=120.0)
    assert res.status == 0, res.describe()
    return (e.download(np.float64), e16.download(np.float32), cnt.download(np.int64), ph.download(np.uint64),
            rh.download(np.uint64), tp.download(np.float64))


@pytest.mark.parametrize("shape", [(1, 1, 10, 10, 300, 4), (2, 2, 16, 16, 2000, 5), (3, 2, 8, 12, 500, 12),
                                   (4, 4, 32, 32, 20000, 6), (8, 8, 128, 128, 50000, 2)])
def test_pic_matches_oracle(gm, orc, shape):
    px, py, nx, ny, np_, cycles = shape
    print(f"[gpu] pic_exact {px}x{py} ranks, {nx}x{ny} cells, {np_} particles/rank, {cycles} cycles", flush=True)
    e, e16, cnt, ph, rh, _ = run_pic(gm, px, py, nx, ny, np_, cycles)
    d = orc.pic_run(orc.pic_params(px, py, nx, ny, np_, cycles, seed=2024))
    assert np.array_equal(e.view(np.uint64), np.asarray(d["energy"], dtype=np.float64).view(np.uint64))
    assert not e16.any()  # pic_exact leaves the BF16-wire energies alone
    assert np.array_equal(cnt, np.asarray(d["counts"], dtype=np.int64))
    assert int(cnt.sum()) == px * py * np_  # migration conserves particles
    assert np.array_equal(ph, np.asarray(d["part_hash"], dtype=np.uint64))
    assert np.array_equal(rh, np.asarray(d["rho_hash"], dtype=np.uint64))


def test_pic_phase_times_are_reported(gm):
    _, _, cnt, _, _, tp = run_pic(gm, 2, 2, 16, 16, 1000, 3, want_phases=True)
    assert int(cnt.sum()) == 4000
    assert tp[6] > 0 and all(t >= 0 for t in tp[:6]) and tp[:6].sum() <= tp[6] * 1.0001


@pytest.mark.parametrize("shape,P", [((2, 2, 16, 16, 500, 5), 3), ((4, 4, 32, 32, 5000, 6), 16),
                                     ((8, 8, 128, 128, 20000, 3), 64)])
def test_particle_decomposition_pic_matches_the_global_oracle(gm, orc, shape, P):
    # The reference's sputniPIC pattern (progs/pic_pdecomp.hip): field Bcast,
    # particle Scatterv / Gatherv with uneven counts, FP32 Reduce of the
    # counts, FP64 Reduce of the energy; against orc_pic_global.
    px, py, nx, ny, np_, cycles = shape
    n = px * py * np_
    print(f"[gpu] pic_pdecomp {shape} P={P}", flush=True)
    f = lambda x: struct.unpack("<I", struct.pack("<f", x))[0]  # noqa: E731
    e = gm.DeviceBuffer(8 * cycles)
    e.zero()
    g = gm.DeviceBuffer(16 * n)
    g.zero()
    res = gm.run("pic_pdecomp", P, [0, px, py, nx, ny, np_, cycles, 2024, f(0.02), f(0.01), f(0.001), e, g, 0],
                 heap_bytes=P * (16 * (n // P + 1) + 24 * px * nx * py * ny + (2 << 20)), timeout_s=120.0)
    assert res.status == 0, res.describe()
    eg, qg = orc.pic_global(orc.pic_params(*shape, seed=2024))
    assert np.array_equal(g.download(np.float32).reshape(-1, 4).view(np.uint32),
                          np.asarray(qg, dtype=np.float32).reshape(-1, 4).view(np.uint32))
    got = e.download(np.float64)
    assert np.max(np.abs(got - eg) / eg) <= 1e-13
