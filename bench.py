#!/usr/bin/env python3
"""Benchmark of the device-resident MPI hot path on one MI355X.

Headline workload (`value`, the p2p configuration): 256 MPI ranks, one
workgroup each, in one persistent kernel; every rank passes its 64 MiB
seeded uint8 buffer to rank+1 mod 256 (MPI_Irecv + MPI_Isend +
MPI_Waitall).  A step is one ring shift of all ranks.  value = whole-job
throughput in GB/s with the algorithmic byte count 2 * P * S per step (each
payload read once and written once), from HIP events around the timed
launch; the in-kernel MPI_Wtime timing (max over ranks) is reported beside
it.  The received buffers are verified against the seeded payloads after
the timed launch.

The other configurations are reported on the same line under "configs",
each with its own CPU baseline:
  pi       cpi.c as a rank program at 64 ranks, N = 1e8: Reduce time, total
           time, error against the FP64 fold and against pi; host baseline:
           the reference's cpi.c under host MPI at 64 ranks;
  p2p      8 B ping-pong of one pair and of 128 concurrent pairs, and the
           8 B - 64 MiB ping-pong sweep (one pair, 128 pairs);
  npb_is   NPB IS class B at 128 ranks: keys/s, Alltoallv GB/s and roofline,
           per-phase times; CPU baseline: the reference's is.c under host MPI
           at 128 ranks (else the oracle's restatement on one host thread);
  allreduce  the exact FP32 Allreduce of 256 ranks x 16 MiB; with --wire
           also the compressed-wire Allreduce (BF16, FP8 e4m3) against the
           exact path at 256 ranks x 1M floats, with their error against FP64;
  pic      the particle-in-cell cycle at 8 x 8 ranks x 128 x 128 cells x
           1M particles per rank, 100 cycles: ms per cycle, phase split,
           communication share, HBM roofline on 48 B per particle per cycle,
           energies against the oracle; CPU baseline: the oracle's cycle.
  autotune the launch-knob autotuner at 256 ranks: chosen knobs, objective
           against the defaults', every trial.
Each side configuration runs in a child process of its own (--config NAME).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dphpc-gpu_mpi_amd"))

METRIC = "Allreduce/Alltoallv GB/s (% HBM peak); p2p latency µs; sort keys/s vs host MPI"
HBM_PEAK_GBPS = 8000.0
SEED = 0x5EED


def cpu_baseline(nbytes, budget_s=20.0):
    """Host MPI (MPICH) ring harness on the host cores; runs before any GPU use."""
    oracle = os.path.join(ROOT, "oracle")
    exe = os.path.join(oracle, "_build", "hostmpi_p2p")
    mpirun = "/opt/conda/bin/mpirun"
    if not os.path.exists(mpirun):
        return None
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", oracle, "_build/hostmpi_p2p"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    cores = min(os.cpu_count() or 1, 16)
    # a short probe sizes the sample to roughly `budget_s` of work
    probe = subprocess.run([mpirun, "-np", str(cores), exe, "ring", str(nbytes), "2", "1", str(SEED)],
                           capture_output=True, text=True, timeout=300)
    p = json.loads(probe.stdout.strip().splitlines()[-1])
    per_iter = max(p["seconds"] / 2, 1e-6)
    iters = int(max(3, min(200, budget_s / per_iter)))
    out = subprocess.run([mpirun, "-np", str(cores), exe, "ring", str(nbytes), str(iters), "1", str(SEED)],
                         capture_output=True, text=True, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    ver = subprocess.run([mpirun, "--version"], capture_output=True, text=True).stdout.split("\n")
    mpi_version = next((l.strip() for l in ver if "Version" in l), "MPICH")
    return {
        "value": round(r["gbps"], 3), "unit": "GB/s", "cores": cores, "kind": "port",
        "sample": (f"host MPI ring shift, {cores} ranks (one per core) x {nbytes} B, {iters} steps, "
                   f"same seeded payload and verification; {mpi_version}; mismatches={r['mismatches']}"),
    }


def measured_traffic(P, S, threads, steps):
    """HBM bytes per timed launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs recorded in profiles/ (per ring step, calibrated against a plain
    copy kernel), when they were taken on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "p2p_ring_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if (rec.get("ranks"), rec.get("message_bytes"), rec.get("threads")) != (P, S, threads):
        return None
    return round(rec["hbm_bytes_per_step"] * steps)


def copy_calibration(nbytes):
    """Two launches of a plain dwordx4 copy of nbytes (read nbytes, write
    nbytes): what FETCH_SIZE / WRITE_SIZE read for it scales the ring's."""
    import gpumpi as gm
    gm.load()
    nbytes = nbytes // 16 * 16
    a = gm.DeviceBuffer(nbytes)
    b = gm.DeviceBuffer(nbytes)
    a.zero()
    for _ in range(2):
        gm.copy_calibration(b, a, nbytes)
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 2}), flush=True)
    return 0


def cpi_host_baseline(P=64, n=100000000, runs=3):
    """The reference's cpi.c (built with host mpicc into oracle/_ref/cpi)
    under host MPI at P ranks; min over `runs` of its printed wall clock
    time.  Runs before the GPU is touched."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cpi")
    mpirun = "/opt/conda/bin/mpirun"
    if not (os.path.exists(exe) and os.path.exists(mpirun)):
        return None
    best = None
    for _ in range(runs):
        out = subprocess.run([mpirun, "-np", str(P), exe, str(n)], capture_output=True, text=True,
                             timeout=300).stdout
        for line in out.splitlines():
            if "wall clock time" in line:
                t = float(line.split("=")[1])
                best = t if best is None else min(best, t)
    if best is None:
        return None
    cores = min(os.cpu_count() or 1, 16)
    return {"value": round(best, 6), "unit": "s (total)", "cores": cores, "kind": "reference",
            "sample": f"toolchain_tests/pi/cpi.c under host MPI, {P} ranks on {cores} cores "
                      f"(oversubscribed {P / cores:.0f}x), N={n}, min of {runs} runs"}


def is_host_baseline():
    """NPB IS class B at 128 ranks on the host: the reference's own is.c
    under host MPI (oracle/_ref/npb/is.B.x, built by `make -C oracle ref`;
    its reported time and Mop/s over the 10 timed iterations, is.c:1148),
    else the oracle's restatement on one host thread."""
    exe = os.path.join(ROOT, "oracle", "_ref", "npb", "is.B.x")
    mpirun = "/opt/conda/bin/mpirun"
    if os.path.exists(exe) and os.path.exists(mpirun):
        out = subprocess.run([mpirun, "-np", "128", exe], capture_output=True, text=True, timeout=600).stdout
        vals = {}
        for line in out.splitlines():
            if "=" in line:
                k, _, v = line.partition("=")
                vals[k.strip()] = v.strip()
        if vals.get("Verification") == "SUCCESSFUL" and "Time in seconds" in vals:
            t = float(vals["Time in seconds"])
            cores = min(os.cpu_count() or 1, 16)
            return {"value": round(10 * (1 << 25) / t / 1e6, 1), "unit": "Mkeys/s", "cores": cores,
                    "kind": "reference",
                    "sample": f"the reference's NPB IS (is.c) class B under host MPI, 128 ranks on {cores} cores "
                              f"(oversubscribed), 10 timed iterations: {t:.2f} s, verification SUCCESSFUL"}
    for p in (os.path.join(ROOT, "oracle"),):
        if p not in sys.path:
            sys.path.insert(0, p)
    import oracle
    oracle.load()
    res, _, _ = oracle.is_run("B", 128)
    return {"value": round(10 * (1 << 25) / res.timed_seconds / 1e6, 1), "unit": "Mkeys/s", "cores": 1,
            "kind": "port",
            "sample": f"NPB IS class B (2^25 keys), 128 simulated ranks, 10 timed iterations: "
                      f"{res.timed_seconds:.2f} s; verification passed={res.passed_verification}"}


def pi_config(gm, np, cpu):
    """cpi.c as a rank program: 64 ranks, N = 1e8, FP64 partials reduced to
    rank 0 (a rank-order fold, bit-equal to the host FP64 fold of the same
    partials)."""
    import math
    P, n = 64, 100000000
    out = gm.DeviceBuffer(8 * (2 + P))
    out.zero()
    rt = gm.DeviceBuffer(8 * P)
    rt.zero()
    gm.run("cpi", P, [0, n, out, rt], timeout_s=120.0)
    got = out.download(np.float64)
    parts = got[2:]
    fold = float(parts[0])
    for x in parts[1:]:
        fold = float(x) + fold
    pi = float(got[0])
    return {"ranks": P, "intervals": n, "reduce_us": round(float(rt.download(np.float64).max()) * 1e6, 3),
            "total_s": round(float(got[1]), 6), "pi": pi,
            "rel_err_vs_fp64_fold": abs(pi - fold) / abs(fold), "abs_err_vs_pi": abs(pi - math.pi),
            "cpu_baseline": cpu}


def p2p_config(gm, np, sb, rb, stride, P):
    """Ping-pong (Send/Recv | Recv/Send, mpi_communication.c:16-51) of one
    pair and of 128 concurrent pairs at 8 B .. 64 MiB; one-way time =
    RTT / 2, GB/s on 2 * S bytes per one-way transfer (per pair)."""
    sizes = [8 << k for k in range(24)]
    iters = [1000 if s <= (64 << 10) else 200 if s <= (1 << 20) else 50 if s <= (16 << 20) else 20 for s in sizes]
    warm = [max(2, i // 10) for i in iters]
    out = {}
    for pairs in (1, 128):
        szb = gm.DeviceBuffer.from_array(np.array(sizes, dtype=np.uint64))
        itb = gm.DeviceBuffer.from_array(np.array(iters, dtype=np.int64))
        wmb = gm.DeviceBuffer.from_array(np.array(warm, dtype=np.int64))
        tb = gm.DeviceBuffer(8 * P * len(sizes))
        tb.zero()
        gm.run("p2p_pingpong_sweep", P, [0, szb, len(sizes), itb, wmb, sb, rb, stride, tb, pairs],
               threads_per_rank=256, timeout_s=300.0)
        t = tb.download(np.float64).reshape(len(sizes), P)
        rows = []
        for k, S in enumerate(sizes):
            rtt = float(t[k, :2 * pairs].max()) / iters[k]
            one_way = rtt / 2
            rows.append({"bytes": S, "one_way_us": round(one_way * 1e6, 3),
                         "gbps_per_pair": round(2 * S / one_way / 1e9, 2),
                         "gbps_all_pairs": round(pairs * 2 * S / one_way / 1e9, 2)})
        out[f"pingpong_{pairs}_pairs"] = rows
    out["latency_us_8B_one_way"] = out["pingpong_1_pairs"][0]["one_way_us"]
    out["latency_us_8B_one_way_128_pairs"] = out["pingpong_128_pairs"][0]["one_way_us"]
    return out


IS_PHASES = ["bucket histogram", "bucket pointers", "scatter to buckets", "Allreduce of bucket sizes",
             "redistribution plan", "Alltoall of counts", "Alltoallv of keys", "rank sums", "key counts",
             "count prefix", "partial verification"]


# The largest user LDS of a 128-rank launch (160 KiB less the runtime's part):
# the widest LDS windows for the edge ranks' key counts.
IS_USER_LDS = 151552


def is_config(gm, np, threads, cpu):
    """NPB IS class B at 128 ranks: keys/s over the 10 timed iterations
    (is.c:1148), its Alltoallv's GB/s on 2 * 2^25 * 4 bytes per iteration,
    and rank 0's time per phase."""
    res = gm.DeviceBuffer(8 * 8)
    res.zero()
    tout = gm.DeviceBuffer(8 * 4)
    ph = gm.DeviceBuffer(8 * len(IS_PHASES))
    ph.zero()
    gm.run("npb_is", 128, [0, ord("B"), res, tout, 0, 10, ph], threads_per_rank=threads,
           user_lds_bytes=IS_USER_LDS, heap_bytes=128 * (32 << 20), max_requests=64, timeout_s=120.0)
    r = res.download(np.int64)
    tt = tout.download(np.float64)
    phases = ph.download(np.float64)
    a2a = 2 * 10 * (1 << 25) * 4 / tt[1] / 1e9
    return {"verified": bool(r[0] == 50 + 128), "ranks": 128, "threads_per_rank": threads,
            "mkeys_per_s": round(10 * (1 << 25) / tt[0] / 1e6, 1), "ms_per_iteration": round(tt[0] / 10 * 1e3, 4),
            "alltoallv_gbps": round(a2a, 1),
            "roofline": {"bound": "hbm", "phase": "Alltoallv of keys", "achieved": round(a2a, 1),
                         "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(a2a / HBM_PEAK_GBPS, 4),
                         "traffic": None},
            "phase_ms_per_iteration_rank0": {name: round(float(v) / 10 * 1e3, 4) for name, v in zip(IS_PHASES, phases)},
            "cpu_baseline": cpu}


def allreduce_config(gm, np, threads, wire=False):
    """Exact FP32 Allreduce of 256 ranks x 16 MiB (algorithmic 2*P*N*4
    bytes), and, with `wire`, at 256 ranks x 1M floats (N(0,1), seeded per
    rank) the exact path against the BF16 and FP8-e4m3 wires: time per call,
    speed-up over exact, and the normwise error of the result against the
    FP64 sum."""
    out = {}
    P, n, iters = 256, 1 << 22, 10
    ib = gm.DeviceBuffer(4 * n * P)
    ob = gm.DeviceBuffer(4 * n * P)
    ib.zero()
    times = gm.DeviceBuffer(8 * P)
    gm.run("coll_reduce", P, [0, gm.DT["MPI_FLOAT"], gm.OP["MPI_SUM"], n, ib, ob, 4 * n, -1, 0, 1, iters, times, 0, 0],
           threads_per_rank=threads, timeout_s=300.0)
    t = float(times.download(np.float64).max()) / iters
    gbps = 2.0 * P * n * 4 / t / 1e9
    out["exact_f32_256x16MiB"] = {"gbps": round(gbps, 1), "ms_per_call": round(t * 1e3, 4),
                                  "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS,
                                               "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4)}}
    del ib, ob
    if not wire:
        out["wire_256x1M_f32"] = {"status": "not run by default (--wire runs it): the compressed-wire Allreduce "
                                            "shares the program of an unexplained GPU memory fault"}
        return out
    n = 1 << 20
    ib = gm.DeviceBuffer(4 * n * P)
    ob = gm.DeviceBuffer(4 * n * P)
    for r in range(P):
        gm.fill_normal(ib, n, 1234, r, offset_elems=r * n)
    x = ib.download(np.float32).reshape(P, n)
    ref = x.astype(np.float64).sum(axis=0)
    del x
    wires = {}
    for wire, name in ((0, "exact"), (1, "bf16"), (2, "fp8_e4m3")):
        ob.zero()
        times = gm.DeviceBuffer(8 * P)
        gm.run("coll_reduce", P, [0, gm.DT["MPI_FLOAT"], gm.OP["MPI_SUM"], n, ib, ob, 4 * n, -1, 0, 1, iters, times,
                                  0, wire], threads_per_rank=threads, timeout_s=300.0)
        tw = float(times.download(np.float64).max()) / iters
        y = ob.download(np.float32, n).astype(np.float64)  # rank 0's result (every rank's is the same)
        err = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
        wires[name] = {"ms_per_call": round(tw * 1e3, 4), "gbps_fp32_bytes": round(2.0 * P * n * 4 / tw / 1e9, 1),
                       "rel_err_vs_fp64": err}
    for name in ("bf16", "fp8_e4m3"):
        wires[name]["speedup_vs_exact"] = round(wires["exact"]["ms_per_call"] / wires[name]["ms_per_call"], 3)
    out["wire_256x1M_f32"] = wires
    return out


def pic_host_baseline():
    """The oracle's restatement of the C5 cycle at its full shape (8 x 8
    ranks, 128 x 128 cells and 1M particles per rank) on one host thread:
    seconds per cycle from a 2-cycle run less the set-up (a 0-cycle run).
    Also returns the two cycles' FP64 energies, which the device's first two
    cycles must equal bit for bit."""
    import time
    for p in (os.path.join(ROOT, "oracle"),):
        if p not in sys.path:
            sys.path.insert(0, p)
    import oracle
    oracle.load()
    t0 = time.perf_counter()
    oracle.pic_run(oracle.pic_params(*PIC_SHAPE[:5], 0))
    t1 = time.perf_counter()
    d = oracle.pic_run(oracle.pic_params(*PIC_SHAPE[:5], 2))
    t2 = time.perf_counter()
    per_cycle = ((t2 - t1) - (t1 - t0)) / 2
    parts = PIC_SHAPE[0] * PIC_SHAPE[1] * PIC_SHAPE[4]
    return {"value": round(per_cycle * 1e3, 2), "unit": "ms per cycle", "cores": 1, "kind": "port",
            "sample": f"the oracle's PIC cycle at the full shape ({parts} particles), 2 cycles less a 0-cycle "
                      f"set-up run: {t2 - t1:.2f} s and {t1 - t0:.2f} s"}, [float(x) for x in d["energy"]]


# C5: px, py, nx, ny, particles per rank, cycles
PIC_SHAPE = (8, 8, 128, 128, 1000000, 100)
PIC_PHASES = ["push and partition", "particle exchange (Sendrecv)", "deposit and energy fold",
              "energy Allreduce", "density halo exchange (Sendrecv)", "field"]
# Bytes each particle must cross HBM per cycle: the push reads and writes it
# (2 x 16 B) and the deposition reads it again (16 B).
PIC_BYTES_PER_PARTICLE = 48


def pic_config(gm, np, threads, cpu):
    """C5 particle-in-cell at 8 x 8 ranks x 128 x 128 cells x 1M particles per
    rank, 100 cycles (`pic_exact`: FP64 energy Allreduce): time per cycle,
    rank 0's phase split, the communication share, and the first two
    cycles' energies against the oracle's."""
    import struct
    px, py, nx, ny, npr, cycles = PIC_SHAPE
    P = px * py
    cap = 2 * npr + 1024
    e = gm.DeviceBuffer(8 * cycles)
    cnt = gm.DeviceBuffer(8 * P)
    tp = gm.DeviceBuffer(8 * 7)
    for b in (e, cnt, tp):
        b.zero()
    f = lambda x: struct.unpack("<I", struct.pack("<f", x))[0]  # noqa: E731
    res = gm.run("pic_exact", P, [0, px, py, nx, ny, npr, cycles, 2024, f(0.02), f(0.01), f(0.001), e, 0, cnt, 0, 0,
                                  tp, cap], threads_per_rank=threads, heap_bytes=P * (49 * cap + (2 << 20)),
                 timeout_s=300.0)
    energy = e.download(np.float64)
    t = tp.download(np.float64)
    counts = cnt.download(np.int64)
    cpu_row, cpu_energy = cpu if cpu else (None, None)
    per_cycle = t[6] / cycles
    comm = t[1] + t[3] + t[4]
    gbps = P * npr * PIC_BYTES_PER_PARTICLE / per_cycle / 1e9
    out = {"ranks": P, "cells_per_rank": nx * ny, "particles_per_rank": npr, "cycles": cycles,
           "threads_per_rank": threads, "status": res.status,
           "particles_conserved": bool(int(counts.sum()) == P * npr),
           "ms_per_cycle": round(per_cycle * 1e3, 4),
           "phase_ms_per_cycle_rank0": {k: round(float(v) / cycles * 1e3, 4) for k, v in zip(PIC_PHASES, t[:6])},
           "communication_share": round(float(comm / t[6]), 4),
           "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                        "frac": round(gbps / HBM_PEAK_GBPS, 4), "traffic": None,
                        "algorithmic_bytes": f"{PIC_BYTES_PER_PARTICLE} B per particle per cycle"},
           "energy_first": float(energy[0]), "energy_last": float(energy[-1]),
           "cpu_baseline": cpu_row}
    if cpu_energy is not None:
        out["energy_bit_equal_to_oracle_first_2_cycles"] = bool(
            np.array_equal(energy[:2].view(np.uint64), np.asarray(cpu_energy, dtype=np.float64).view(np.uint64)))
    return out


def autotune_config(gm):
    """The autotuner (gpumpi/tune.py) at 256 ranks: one coordinate-descent
    pass over threads per rank, copy-split threshold, eager threshold and
    channel slots, on the ring's geometric-mean step time over 8 B .. 16 MiB.
    Reports the choices, the objective against the defaults', and every
    trial."""
    from gpumpi import tune
    res = tune.tune_device(gm, 256, repeats=1, log=lambda m: print(f"[tune] {m}", file=sys.stderr, flush=True))
    return {"ranks": 256, "sizes": res["sizes"], "chosen": res["best"],
            "objective_us": round(res["objective"] * 1e6, 3),
            "default_objective_us": round(res["default_objective"] * 1e6, 3),
            "speedup_vs_default": round(res["speedup_vs_default"], 4), "seconds": res["seconds"],
            "trials": [{"knobs": t["knobs"], "objective_us": None if t["objective"] is None
                        else round(t["objective"] * 1e6, 3)} for t in res["trials"]]}


CONFIGS = ("pi", "p2p", "npb_is", "allreduce", "pic", "autotune")


def run_config(name, args):
    """One side configuration in this process: its host baseline first,
    then the GPU.  Returns the configuration's JSON object."""
    cpu = None
    if not args.no_cpu_baseline:
        cpu = {"pi": cpi_host_baseline, "npb_is": is_host_baseline, "pic": pic_host_baseline}.get(name,
                                                                                                   lambda: None)()
    import numpy as np
    import gpumpi as gm
    gm.load()
    try:
        if name == "pi":
            return pi_config(gm, np, cpu)
        if name == "p2p":
            P, S = 256, 64 << 20
            sb = gm.DeviceBuffer(S * P)
            rb = gm.DeviceBuffer(S * P)
            gm.fill_payload(sb, S, S, SEED, 0, P)
            return p2p_config(gm, np, sb, rb, S, P)
        if name == "npb_is":
            return is_config(gm, np, args.threads, cpu)
        if name == "allreduce":
            return allreduce_config(gm, np, args.threads, args.wire)
        if name == "pic":
            return pic_config(gm, np, args.threads, cpu)
        if name == "autotune":
            return autotune_config(gm)
        raise ValueError(name)
    finally:
        gm.release_world()


def side_configs(args):
    """Every side configuration in a child process of its own (this script
    with --config NAME), so that a failure in one is reported in its entry
    and leaves the others and the headline line intact."""
    out = {}
    for name in CONFIGS:
        print(f"[bench] config {name}", file=sys.stderr, flush=True)
        cmd = [sys.executable, os.path.abspath(__file__), "--config", name, "--threads", str(args.threads)]
        cmd += ["--no-cpu-baseline"] if args.no_cpu_baseline else []
        cmd += ["--wire"] if args.wire else []
        try:
            p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True, timeout=900)
        except subprocess.TimeoutExpired:
            out[name] = {"status": "failed: no result within 900 s"}
            continue
        lines = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode == 0 and lines:
            out[name] = json.loads(lines[-1])
        else:
            out[name] = {"status": f"failed: exit status {p.returncode}", "output_tail": p.stdout[-400:]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ranks", type=int, default=256)
    ap.add_argument("--bytes", type=int, default=64 << 20)
    ap.add_argument("--threads", type=int, default=512,
                    help="threads per rank workgroup (512: eight waves, one rank per CU at 256 ranks)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the Allreduce / NPB IS side measurements")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel over BYTES (reference for HBM counters) and exit")
    ap.add_argument("--wire", action="store_true",
                    help="also time the BF16 / FP8 compressed-wire Allreduce against the exact path")
    ap.add_argument("--config", choices=CONFIGS, help="run one side configuration alone and print its JSON")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the whole MPI world lives in one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.config:
        print(json.dumps(run_config(args.config, args)), flush=True)
        return 0

    # The host baseline runs before the GPU is touched.
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.bytes)

    import numpy as np
    import gpumpi as gm
    gm.load()
    P, S = args.ranks, args.bytes
    stride = (S + 4095) // 4096 * 4096
    sb = gm.DeviceBuffer(stride * P)
    rb = gm.DeviceBuffer(stride * P)
    rb.zero()
    gm.fill_payload(sb, stride, S, SEED, 0, P)

    def ring(program, iters):
        sz = gm.DeviceBuffer.from_array(np.array([S], dtype=np.uint64))
        it = gm.DeviceBuffer.from_array(np.array([iters], dtype=np.int64))
        wm = gm.DeviceBuffer.from_array(np.array([0], dtype=np.int64))
        times = gm.DeviceBuffer(8 * P)
        res = gm.run(program, P, [0, sz, 1, it, wm, sb, rb, stride, times],
                     threads_per_rank=args.threads, timeout_s=600.0)
        return res, times.download(np.float64)

    # Warm-up launch under its own kernel name (p2p_ring_warmup), then the
    # timed launch of exactly K steps (p2p_ring_sweep): a kernel trace of this
    # command lists the timed launch alone under p2p_ring_sweep.
    ring("p2p_ring_warmup", args.warmup)
    res, t = ring("p2p_ring_sweep", args.steps)
    kernel_s = res.kernel_ms / 1e3
    bytes_per_step = 2.0 * P * S
    gbps = bytes_per_step * args.steps / kernel_s / 1e9
    wtime_gbps = bytes_per_step * args.steps / float(t.max()) / 1e9
    bad = gm.verify_payload(rb, stride, S, SEED, -1, P, P)

    # in-GPU ping-pong latency, 8 B, pair (0, 1) inside the same 256-rank world,
    # with four waves per rank (idle helper waves poll LDS beside wave 0, so
    # fewer of them leave the small-message path less contended)
    ppb = gm.DeviceBuffer(4096 * P)
    sz = gm.DeviceBuffer.from_array(np.array([8], dtype=np.uint64))
    it = gm.DeviceBuffer.from_array(np.array([1000], dtype=np.int64))
    wm = gm.DeviceBuffer.from_array(np.array([100], dtype=np.int64))
    pt = gm.DeviceBuffer(8 * P)
    lat_threads = 256
    gm.run("p2p_pingpong_sweep", P, [0, sz, 1, it, wm, ppb, ppb, 4096, pt, 1],
           threads_per_rank=lat_threads, timeout_s=120.0)
    rtt_us = float(pt.download(np.float64)[:2].max()) / 1000 * 1e6

    line = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res.kernel_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic: seeded splitmix64 uint8 payload, verified after the timed launch",
        "config": {"workload": f"p2p ring shift: {P} ranks (one workgroup each, {args.threads} threads) "
                               f"x {S} B per rank per step, MPI_Irecv+MPI_Isend+MPI_Waitall",
                   "ranks": P, "message_bytes": S, "parallelism": "1 GPU, 1 persistent kernel"},
        "roofline": {"bound": "hbm", "achieved": round(gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(gbps / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(P, S, args.threads, args.steps)},
        "cpu_baseline": cpu,
        "p2p_latency_us_8B_one_way": round(rtt_us / 2, 3),
        "p2p_rtt_us_8B": round(rtt_us, 3),
        "p2p_latency_threads_per_rank": lat_threads,
        "ring_gbps_mpi_wtime": round(wtime_gbps, 2),
        "payload_mismatches": bad,
    }
    del sb, rb, ppb
    gm.release_world()
    if not args.no_extras:
        line["configs"] = side_configs(args)
        line["configs"]["batched_allreduce"] = {
            "status": "not measured: the batched uneven Allreduce at its 1024-rank shape is kept off the GPU "
                      "(unexplained GPU memory faults of that family); its parity runs at 96 ranks in the GPU tests"}
    print(json.dumps(line), flush=True)
    return 0 if bad == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
