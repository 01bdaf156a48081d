"""Autotuner for the launch knobs of the device MPI runtime.

It picks, for a world of P ranks on this device:
  threads_per_rank  waves per rank, and so how many ranks share a CU;
  helper_min_bytes  the payload size from which a copy is split across the
                    rank's waves (the copy chunking);
  eager_bytes       the eager threshold: messages up to it travel inside a
                    channel slot, larger ones by a rendezvous direct copy;
  chan_slots        slots per (source, destination) channel.

The objective is the ring shift of the p2p workload (every rank Irecv from
rank-1, Isend to rank+1, Waitall) over a ladder of message sizes from 8 B to
16 MiB: the geometric mean over sizes of the time per step (max over ranks),
so that latency-bound and bandwidth-bound sizes weigh alike.  The search is
one pass of coordinate descent from the runtime's defaults, each knob over
its candidates with the others held; a candidate the launcher refuses (a
world that cannot be co-resident, say) is skipped.  Each candidate is timed
`repeats` times and the fastest run counts.

The choices go to a tuning file of "key value" lines, which
gpumpi_apply_tuning (include/gpumpi/host.h), `---gpumpi --tuning FILE` and
gpumpi.default_config(P, tuning=FILE) apply.

    python -m gpumpi.tune --ranks 256 -o gpumpi.tuning
"""
import argparse
import json
import os
import sys
import time

KNOBS = (
    ("threads_per_rank", (256, 512, 1024)),
    ("helper_min_bytes", (4096, 16384, 65536, 262144)),
    ("eager_bytes", (192, 448, 960, 1984, 4032)),
    ("chan_slots", (4, 8, 16)),
)
SIZES = (8, 512, 8192, 131072, 2 << 20, 16 << 20)
TUNED_KEYS = tuple(k for k, _ in KNOBS)


def ring_iters(size):
    """Timed steps per size: about the same time at every size."""
    return 200 if size <= 8192 else 60 if size <= 131072 else 12


def ring_objective(gm, nranks, knobs, sizes=SIZES, stride=None):
    """Geometric mean over `sizes` of the ring's seconds per step (max over
    ranks) under launch knobs `knobs`.  Raises gm.GpumpiError if the launch
    is refused or fails."""
    import numpy as np
    stride = stride or max(sizes)
    key = ("_tune_buffers", nranks, stride)
    bufs = getattr(ring_objective, "_bufs", None)
    if bufs is None or bufs[0] != key:
        sb = gm.DeviceBuffer(stride * nranks)
        rb = gm.DeviceBuffer(stride * nranks)
        gm.fill_payload(sb, stride, stride, 0x5EED, 0, nranks)
        ring_objective._bufs = bufs = (key, sb, rb)
    _, sb, rb = bufs
    iters = [ring_iters(s) for s in sizes]
    szb = gm.DeviceBuffer.from_array(np.array(sizes, dtype=np.uint64))
    itb = gm.DeviceBuffer.from_array(np.array(iters, dtype=np.int64))
    wmb = gm.DeviceBuffer.from_array(np.array([max(2, i // 10) for i in iters], dtype=np.int64))
    tb = gm.DeviceBuffer(8 * nranks * len(sizes))
    tb.zero()
    gm.run("p2p_ring_sweep", nranks, [0, szb, len(sizes), itb, wmb, sb, rb, stride, tb], timeout_s=120.0, **knobs)
    t = tb.download(np.float64).reshape(len(sizes), nranks).max(axis=1) / np.array(iters, dtype=np.float64)
    if not np.all(t > 0):
        raise gm.GpumpiError("ring timing missing")
    return float(np.exp(np.log(t).mean()))


def autotune(measure, defaults, knobs=KNOBS, repeats=2, log=None):
    """Coordinate descent over `knobs` from `defaults` (a dict of the tuned
    keys).  measure(dict) -> objective (lower is better), raising for a
    refused candidate.  Returns {"best": dict, "objective": float,
    "default_objective": float, "trials": [...]}."""
    def timed(cfg):
        best = None
        for _ in range(repeats):
            v = measure(dict(cfg))
            best = v if best is None else min(best, v)
        return best

    cur = {k: defaults[k] for k, _ in knobs}
    trials = []
    base = timed(cur)
    trials.append({"knobs": dict(cur), "objective": base})
    best_val = base
    for key, cands in knobs:
        start = cur[key]  # the point this sweep starts from is measured already
        for v in cands:
            if v == start:
                continue
            cand = dict(cur, **{key: v})
            try:
                val = timed(cand)
            except Exception as e:  # refused or failed launch: not a candidate
                trials.append({"knobs": cand, "objective": None, "error": str(e)[:200]})
                if log:
                    log(f"{key}={v}: skipped ({str(e)[:80]})")
                continue
            trials.append({"knobs": cand, "objective": val})
            if log:
                log(f"{key}={v}: {val:.6g}")
            if val < best_val:
                best_val, cur = val, cand
    return {"best": cur, "objective": best_val, "default_objective": base, "trials": trials}


def write_tuning(path, values, meta=None):
    """The tuning file gpumpi_apply_tuning reads."""
    lines = ["# gpumpi tuning: launch knobs chosen by gpumpi.tune"]
    for k in TUNED_KEYS:
        if k in values:
            lines.append(f"{k} {int(values[k])}")
    for k, v in (meta or {}).items():
        lines.append(f"# {k}: {v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_tuning(path):
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].split()
            if len(line) == 2 and line[0] in TUNED_KEYS:
                out[line[0]] = int(line[1])
    return out


def tune_device(gm, nranks, repeats=2, log=None, knobs=KNOBS, sizes=SIZES):
    """autotune with the ring objective on the device."""
    d = gm.default_config(nranks)
    defaults = {k: int(getattr(d, k)) for k in TUNED_KEYS}
    t0 = time.time()
    res = autotune(lambda kn: ring_objective(gm, nranks, kn, sizes), defaults, knobs, repeats, log)
    res["seconds"] = round(time.time() - t0, 2)
    res["ranks"] = nranks
    res["sizes"] = list(sizes)
    res["speedup_vs_default"] = res["default_objective"] / res["objective"]
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="Tune the launch knobs of the device MPI runtime")
    ap.add_argument("--ranks", type=int, default=256)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("-o", "--output", default="gpumpi.tuning")
    a = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import gpumpi as gm
    gm.load()
    res = tune_device(gm, a.ranks, a.repeats, log=lambda m: print(f"[tune] {m}", file=sys.stderr, flush=True))
    write_tuning(a.output, res["best"], {"ranks": a.ranks, "objective_s": res["objective"],
                                         "default_objective_s": res["default_objective"]})
    print(json.dumps({k: res[k] for k in ("best", "objective", "default_objective", "speedup_vs_default",
                                           "seconds")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
