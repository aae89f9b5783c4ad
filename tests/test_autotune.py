"""The autotuner's form selection: tuned.json entries with and without a
chosen form, the parameters a form maps to, and (GPU) a small tuning run
whose chosen form reproduces the Fibonacci recurrence."""
from __future__ import annotations

import json

import numpy as np
import pytest

from loom_amd import abi, autotune, workloads


def test_best_form_defaults_to_hybrid(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "TUNED", tmp_path / "tuned.json")
    assert autotune.best_form("fib_u64") == ("hybrid", autotune.DEFAULT)
    (tmp_path / "tuned.json").write_text(json.dumps(
        {"fib_u64": {"params": {"live_threshold": 0, "waves_per_cu": 16}}}))
    mode, params = autotune.best_form("fib_u64")
    assert mode == "hybrid" and params["live_threshold"] == 0 and params["waves_per_cu"] == 16


def test_best_form_reads_compact(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "TUNED", tmp_path / "tuned.json")
    (tmp_path / "tuned.json").write_text(json.dumps(
        {"fib_u64": {"params": {"live_threshold": 0},
                     "form": {"mode": "compact", "params": {"reconv": 2}}}}))
    mode, params = autotune.best_form("fib_u64")
    assert (mode, params) == ("compact", {"reconv": 2})
    m, p = autotune.params_of_form(mode, params)
    assert m == abi.MODE_COMPACT and p.reconv == 2 and p.live_threshold == 0
    # the hybrid settings stay available for batched launches
    assert autotune.best_params("fib_u64")["live_threshold"] == 0


def test_committed_forms_are_valid():
    data = json.loads(autotune.TUNED.read_text())
    for name, entry in data.items():
        if "form" not in entry:
            continue
        f = entry["form"]
        assert f["mode"] in ("hybrid", "compact"), name
        keys = (set(autotune.SPACE) | set(autotune.WGR_SPACE) if f["mode"] == "hybrid"
                else set(autotune.COMPACT_SPACE))
        assert set(f["params"]) <= keys, name
        # the chosen form is the fastest that keeps the active-lane floor
        assert f["ms"] <= entry["ms"] + 1e-9 or f["active_lane_fraction"] < entry["min_active"]


@pytest.mark.gpu
def test_gpu_tune_small(gpu, tmp_path):
    import loom_amd
    x = workloads.fib_inputs(1 << 14)
    space = {"live_threshold": [0, 48], "regroup": [False], "reconv": [2], "waves_per_cu": [0]}
    entry = autotune.tune("fib_u64", [x], space=space, verbose=False, min_active=0.0)
    assert entry["candidates"] == (2 + len(autotune.COMPACT_SPACE["reconv"])
                                   + len(autotune.WGR_SPACE["regroup_phase"])
                                   * len(autotune.WGR_SPACE["regroup_waves"]))
    f = entry["form"]
    assert f["ms"] == min(f["ms"], entry["ms"])
    m, p = autotune.params_of_form(f["mode"], f["params"])
    k = loom_amd.kernel("fib_u64")
    din = [loom_amd.DeviceArray.from_numpy(x)]
    dout = [loom_amd.DeviceArray(len(x), np.uint64)]
    loom_amd.run(k, m, din, dout, params=p)
    fib = [0, 1]
    while len(fib) < 100:
        fib.append((fib[-1] + fib[-2]) % (1 << 64))
    want = np.array([fib[int(v)] for v in x], dtype=np.uint64)
    assert np.array_equal(dout[0].numpy(), want)
