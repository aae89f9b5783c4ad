"""Autotuner: picks the execution form of a kernel (the hybrid form, the
hybrid form regrouping within workgroups, or the predicated compact form)
and its settings by timing candidates on a sample of the real workload.
Hybrid candidates vary the live-lane refill threshold, the regroup switch,
the reconvergence policy and the persistent grid (waves per CU);
workgroup-regrouping candidates the phase length; compact candidates the
reconvergence policy.

    python -m loom_amd.autotune smallpt9_f32      # tunes, writes tuned.json
    python -m loom_amd.autotune batch:pt_int_mix  # the 64-job mixed batch

Results are stored per kernel in loom_amd/tuned.json: `best_form` gives the
fastest form that keeps the required share of lanes active, `best_params`
the best hybrid settings (batched launches and the serving loop run the
hybrid form), or the defaults for a kernel never tuned.
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

# the compact form: one wave per 64 threads, reconvergence policy only
COMPACT_SPACE = {"reconv": [0, 1, 2]}

# the hybrid form regrouping within workgroups (sweep policy): the phase
# length between two sorts and the workgroup size in waves
WGR_SPACE = {"regroup_phase": [64, 256, 1024], "regroup_waves": [16, 8]}


def best_params(kernel_name: str) -> dict:
    """Tuned hybrid parameters of `kernel_name` (defaults if never tuned)."""
    if TUNED.exists():
        data = json.loads(TUNED.read_text())
        if kernel_name in data:
            return {**DEFAULT, **data[kernel_name]["params"]}
    return dict(DEFAULT)


def best_form(kernel_name: str) -> tuple[str, dict]:
    """The tuned form of `kernel_name`, ("hybrid" | "compact", keyword
    arguments of make_params); the hybrid form with its best settings when
    the kernel was tuned without compact candidates or never tuned."""
    if TUNED.exists():
        data = json.loads(TUNED.read_text())
        form = data.get(kernel_name, {}).get("form")
        if form:
            return form["mode"], dict(form["params"])
    return "hybrid", best_params(kernel_name)


def params_of_form(mode: str, params: dict):
    """(abi mode, loom_params) of a form returned by best_form."""
    import loom_amd
    from loom_amd import abi
    m = abi.MODE_HYBRID if mode == "hybrid" else abi.MODE_COMPACT
    return m, loom_amd.make_params(mode=m, **params)


def tune(kernel_name: str, inputs, reps: int = 2, space: dict | None = None,
         verbose: bool = True, min_active: float = MIN_ACTIVE,
         compact_space: dict | None = None, wgr_space: dict | None = None) -> dict:
    """Time every candidate of `space` (hybrid), `wgr_space` (hybrid,
    regrouping within workgroups) and `compact_space` (compact) on `inputs`
    (numpy columns, copied to HBM once).  The result holds the fastest
    hybrid parameters without workgroup regrouping (what batched launches
    and the serving loop can use) and the fastest form overall, each among
    the candidates whose active-lane fraction is at least `min_active` (the
    fastest of all if none reaches it)."""
    import loom_amd
    from loom_amd import abi

    loom_amd.init(0)
    k = loom_amd.kernel(kernel_name)
    n = len(inputs[0])
    din = [loom_amd.DeviceArray.from_numpy(a) for a in inputs]
    dout = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
    space = space or SPACE
    compact_space = COMPACT_SPACE if compact_space is None else compact_space
    keys = list(space)
    cands = [("hybrid", {**DEFAULT, **dict(zip(keys, combo))})
             for combo in itertools.product(*(space[k2] for k2 in keys))]
    ckeys = list(compact_space)
    cands += [("compact", dict(zip(ckeys, combo)))
              for combo in itertools.product(*(compact_space[k2] for k2 in ckeys))]
    wgr_space = WGR_SPACE if wgr_space is None else wgr_space
    wkeys = list(wgr_space)
    cands += [("hybrid", {**DEFAULT, "regroup": 2, "reconv": 2, **dict(zip(wkeys, combo))})
              for combo in itertools.product(*(wgr_space[k2] for k2 in wkeys))]
    results = []
    for mode, cand in cands:
        m, p = params_of_form(mode, {**cand, "stats": True})
        act = loom_amd.run(k, m, din, dout, params=p).active_lane_fraction  # warm
        # timed without statistics, the form as it is used; short runs are
        # repeated until about 20 ms have passed
        m, p = params_of_form(mode, {**cand, "stats": False})
        t0 = time.perf_counter()
        loom_amd.run(k, m, din, dout, params=p)
        first = (time.perf_counter() - t0) * 1e3
        count = max(reps, min(50, int(20.0 / max(first, 1e-3))))
        t0 = time.perf_counter()
        for _ in range(count):
            loom_amd.run(k, m, din, dout, params=p)
        ms = (time.perf_counter() - t0) * 1e3 / count
        results.append((ms, act, mode, cand))
        if verbose:
            print(f"tune {kernel_name} {mode} {cand} {ms:.3f} ms active {act:.3f}", flush=True)
    results.sort(key=lambda r: r[0])

    def pick(rs):
        ok = [r for r in rs if r[1] >= min_active] or rs
        return ok[0]

    ms, act, _, best = pick([r for r in results if r[2] == "hybrid" and r[3]["regroup"] != 2])
    fms, fact, fmode, fbest = pick(results)
    if fmode == "compact":
        fparams = {k2: fbest[k2] for k2 in ckeys}
    elif fbest["regroup"] == 2:
        fparams = {"regroup": 2, "reconv": 2, **{k2: fbest[k2] for k2 in wkeys}}
    else:
        fparams = {k2: fbest[k2] for k2 in keys}
    return {"params": {k2: best[k2] for k2 in keys}, "ms": ms, "active_lane_fraction": act,
            "form": {"mode": fmode, "params": fparams, "ms": fms,
                     "active_lane_fraction": fact},
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
