"""Work counters of a run (loom_run_stats.counters) against what the
threads did, independently of the schedule: the total of thread-block
steps has a closed form for Collatz (START, three blocks per step, the exit
block), and every schedule-independent counter (thread steps, active ops,
active flops) is equal across forms and policies."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests.test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

FORMS = [
    ("compact_sm", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("compact_topo", abi.MODE_COMPACT, dict(reconv=abi.RECONV_TOPO)),
    ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
    ("hybrid_sweep", abi.MODE_HYBRID, {}),
    ("hybrid_sm_regroup", abi.MODE_HYBRID, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR,
                                                regroup=True, live_threshold=40)),
    ("hybrid_topo_one_wave_per_cu", abi.MODE_HYBRID, dict(reconv=abi.RECONV_TOPO,
                                                          waves_per_cu=1)),
]
SCHEDULE_FREE = (abi.C_THREAD_STEPS, abi.C_ACTIVE_OPS, abi.C_ACTIVE_FLOPS)


def run(name, ins, mode, params):
    print(f"[gpu] counters {name} mode={mode} {params} n={len(ins[0])}", flush=True)
    r = loom_amd.run(name, mode, ins, params=loom_amd.make_params(mode=mode, **params))
    assert r.errors == 0
    return r


def test_collatz_thread_steps_closed_form(gpu):
    x = workloads.collatz_inputs(1 << 20)
    for fname, mode, params in FORMS:
        r = run("collatz_u64", [x], mode, params)
        steps = r.outputs[0].astype(np.uint64)
        assert r.counters[abi.C_THREAD_STEPS] == int(np.sum(2 + 3 * steps)), fname
        # slots cover the active work; dispatches cover every block run
        assert r.counters[abi.C_SLOT_OPS] >= r.counters[abi.C_ACTIVE_OPS] > 0
        assert 64 * r.counters[abi.C_DISPATCHES] >= r.counters[abi.C_THREAD_STEPS]


@pytest.mark.parametrize("name,n", [("smallpt9_f32", 40 * 30 * 8), ("fib_u64", 1 << 16),
                                    ("cubic_bisect_f64", 1 << 14)])
def test_schedule_free_counters_agree(gpu, name, n):
    ins = CASES[name](n)
    ref = None
    for fname, mode, params in FORMS:
        c = run(name, ins, mode, params).counters
        got = [c[i] for i in SCHEDULE_FREE]
        assert all(v > 0 for v in got[:2]), fname
        if ref is None:
            ref = got
        assert got == ref, (name, fname)


def _work(key):
    import json
    from pathlib import Path
    p = Path(__file__).resolve().parent / "golden" / "work_totals.json"
    return json.loads(p.read_text()).get(key)


WORK_FORMS = [
    ("hybrid_tuned", abi.MODE_HYBRID, None),
    ("compact_sm", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
]


@pytest.mark.parametrize("key,name", [("fib_u64_c1", "fib_u64"), ("collatz_u64_c2", "collatz_u64"),
                                      ("smallpt9_c3_first_1048576", "smallpt9_f32"),
                                      ("smallpt9_c3", "smallpt9_f32")])
def test_counters_equal_work_totals(gpu, key, name):
    """The device counters of whole configurations (the numerators of the
    benchmark's active-lane fraction and roofline) equal the committed
    totals: closed forms for fib and Collatz, the oracle's count for the
    path tracer."""
    from loom_amd import autotune
    f = _work(key)
    if f is None:
        pytest.skip(f"no {key} totals in the fixture")
    n = f["threads"]
    if name == "fib_u64":
        ins = [workloads.fib_inputs(n)]
    elif name == "collatz_u64":
        ins = [workloads.collatz_inputs(n)]
    else:
        ins = [a[:n] for a in workloads.pt_inputs(1024, 768, 64)]
    dev = [loom_amd.DeviceArray.from_numpy(a) for a in ins]
    k = loom_amd.kernel(name)
    for fname, mode, params in WORK_FORMS:
        p = autotune.best_params(name) if params is None else params
        print(f"[gpu] work totals {key} {fname} n={n}", flush=True)
        outs = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
        r = loom_amd.run(k, mode, dev, outs, params=loom_amd.make_params(mode=mode, **p))
        assert r.errors == 0
        c = r.counters
        assert (c[abi.C_THREAD_STEPS], c[abi.C_ACTIVE_OPS], c[abi.C_ACTIVE_FLOPS]) == \
            (f["steps"], f["ops"], f["flops"]), (key, fname)
