"""Autotuner for the hybrid backend: picks the live-lane refill threshold,
the regroup switch, the reconvergence policy and the persistent grid
(waves per CU) by timing candidates on a sample of the real workload.

    python -m loom_amd.autotune smallpt9_f32      # tunes, writes tuned.json
    python -m loom_amd.autotune batch:pt_int_mix  # the 64-job mixed batch

Results are stored per kernel in loom_amd/tuned.json; `best_params` reads
them (or returns the defaults) so benchmarks and batched launches start from
measured settings.
"""
from __future__ import annotations

import itertools
import json
import sys
import time
from pathlib import Path

import numpy as np

TUNED = Path(__file__).resolve().parent / "tuned.json"

# the selection keeps at least this fraction of lanes active (the target
# the hybrid form is held to), trading a few percent of speed when needed
MIN_ACTIVE = 0.70

DEFAULT = {"live_threshold": 48, "regroup": False, "reconv": 2, "waves_per_cu": 0,
           "threshold": 0, "refill": True}

SPACE = {
    # 0: refill only once every lane of the wave halted
    "live_threshold": [0, 32, 40, 48, 56],
    "regroup": [False, True],
    "reconv": [0, 1, 2],
    "waves_per_cu": [0, 8, 16, 20],
}


def best_params(kernel_name: str) -> dict:
    """Tuned hybrid parameters of `kernel_name` (defaults if never tuned)."""
    if TUNED.exists():
        data = json.loads(TUNED.read_text())
        if kernel_name in data:
            return {**DEFAULT, **data[kernel_name]["params"]}
    return dict(DEFAULT)


def tune(kernel_name: str, inputs, reps: int = 2, space: dict | None = None,
         verbose: bool = True, min_active: float = MIN_ACTIVE) -> dict:
    """Time every candidate of `space` on `inputs` (numpy columns, copied to
    HBM once) and return the fastest parameters whose active-lane fraction
    is at least `min_active` (the fastest overall if none reaches it)."""
    import loom_amd
    from loom_amd import abi

    loom_amd.init(0)
    k = loom_amd.kernel(kernel_name)
    n = len(inputs[0])
    din = [loom_amd.DeviceArray.from_numpy(a) for a in inputs]
    dout = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
    space = space or SPACE
    keys = list(space)
    results = []
    for combo in itertools.product(*(space[k2] for k2 in keys)):
        cand = {**DEFAULT, **dict(zip(keys, combo))}
        p = loom_amd.make_params(mode=abi.MODE_HYBRID, stats=True, **cand)
        act = loom_amd.run(k, abi.MODE_HYBRID, din, dout, params=p).active_lane_fraction  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            loom_amd.run(k, abi.MODE_HYBRID, din, dout, params=p)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        results.append((ms, act, cand))
        if verbose:
            print(f"tune {kernel_name} {cand} {ms:.3f} ms active {act:.3f}", flush=True)
    results.sort(key=lambda r: r[0])
    ok = [r for r in results if r[1] >= min_active] or results
    ms, act, best = ok[0]
    return {"params": {k2: best[k2] for k2 in keys}, "ms": ms, "active_lane_fraction": act,
            "min_active": min_active, "fastest_any_ms": results[0][0], "n": n,
            "candidates": len(results)}


BATCH_SPACE = {
    "live_threshold": [0, 32, 48, 56],
    "waves_per_cu": [0, 12, 16],
    "reconv": [1, 2],
}


def tune_batch(mixed, reps: int = 2, space: dict | None = None, verbose: bool = True) -> dict:
    """Time every candidate of `space` on a resident mixed batch (one
    persistent launch per run) and return the best parameters."""
    import loom_amd
    from loom_amd import abi

    space = space or BATCH_SPACE
    keys = list(space)
    b = mixed.batch()
    results = []
    for combo in itertools.product(*(space[k2] for k2 in keys)):
        cand = {**DEFAULT, **dict(zip(keys, combo))}
        cand.pop("regroup", None)
        p = loom_amd.make_params(mode=abi.MODE_HYBRID, stats=False, **cand)
        b.run(p)  # warm
        ms = 0.0
        for _ in range(reps):
            ms += b.run(p).ms_main
        ms /= reps
        results.append((ms, cand))
        if verbose:
            print(f"tune batch {cand} {ms:.3f} ms", flush=True)
    results.sort(key=lambda r: r[0])
    return {"params": {k2: results[0][1][k2] for k2 in keys}, "ms": results[0][0],
            "threads": mixed.threads, "candidates": len(results)}


def save(kernel_name: str, entry: dict, path: Path | None = None) -> None:
    path = path or TUNED
    data = json.loads(path.read_text()) if path.exists() else {}
    data[kernel_name] = entry
    path.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")


def main(argv=None) -> int:
    import argparse
    from loom_amd import workloads
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*", default=["smallpt9_f32"])
    ap.add_argument("--out", type=Path, default=TUNED, help="tuned settings file")
    a = ap.parse_args(argv)
    for name in a.names:
        if name.startswith("batch:"):
            from loom_amd import mixed
            entry = tune_batch(mixed.MixedBatch())
        elif name.startswith("smallpt9"):
            # the full frame; the path tracer is held to the active-lane floor
            entry = tune(name, workloads.pt_inputs(1024, 768, 64))
        elif name.startswith("collatz"):
            entry = tune(name, [workloads.collatz_inputs(1 << 24)], min_active=0.0)
        else:
            entry = tune(name, [workloads.fib_inputs(1 << 20)], min_active=0.0)
        save(name, entry, a.out)
        print(json.dumps({name: entry}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
