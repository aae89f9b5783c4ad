"""GPU parity: every built-in kernel, compact and hybrid forms, bit-exact
against the CPU oracle on the same seeded inputs; full-size configs checked
against closed forms and across forms.  Calls go through the C ABI."""
import math

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

MODES = [
    ("compact", abi.MODE_COMPACT, {}),
    ("compact_topo", abi.MODE_COMPACT, dict(reconv=abi.RECONV_TOPO)),
    ("hybrid_t0", abi.MODE_HYBRID, dict(threshold=0, live_threshold=0, regroup=True,
                                        reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("hybrid_topo_live48", abi.MODE_HYBRID, dict(live_threshold=48, regroup=True)),
    ("hybrid_sm_live48", abi.MODE_HYBRID, dict(live_threshold=48, regroup=True,
                                               reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("hybrid_live60_noregroup", abi.MODE_HYBRID, dict(live_threshold=60, regroup=False)),
    ("hybrid_default", abi.MODE_HYBRID, {}),
    ("hybrid_t56_norefill", abi.MODE_HYBRID, dict(threshold=56, live_threshold=0, refill=False,
                                                  regroup=True,
                                                  reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("hybrid_t48_smallq", abi.MODE_HYBRID, dict(threshold=48, live_threshold=40, regroup=True,
                                                queue_capacity=64, waves_per_cu=1,
                                                reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("hybrid_topo_smallq", abi.MODE_HYBRID, dict(live_threshold=40, regroup=True,
                                                 queue_capacity=64, waves_per_cu=1)),
    ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
    ("hybrid_sweep_live48", abi.MODE_HYBRID, dict(live_threshold=48,
                                                  reconv=abi.RECONV_SWEEP)),
    ("hybrid_sweep_smallq", abi.MODE_HYBRID, dict(live_threshold=40, regroup=True,
                                                  queue_capacity=64, waves_per_cu=1,
                                                  reconv=abi.RECONV_SWEEP)),
]
STEP_BUDGET = 1 << 24  # every kernel here finishes far below it


def gpu_run(name, mode, params, inputs):
    p = loom_amd.make_params(max_steps=STEP_BUDGET, mode=mode, **params)
    print(f"[gpu] {name} mode={mode} params={params} n={len(inputs[0])}", flush=True)
    r = loom_amd.run(name, mode, inputs, params=p)
    assert r.errors == 0
    return r


def same(a, b):
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint32 if a.itemsize == 4 else np.uint64),
                              b.view(np.uint32 if b.itemsize == 4 else np.uint64))
    return np.array_equal(a, b)


def int_inputs(name, n, seed):
    k = loom_amd.kernel(name)
    cols = []
    for j, dt in enumerate(k.inputs):
        raw = workloads.splitmix64(seed + j, n)
        bits = np.dtype(dt).itemsize * 8
        cols.append((raw & np.uint64((1 << bits) - 1)).astype(dt))
    return cols


CASES = {
    "fib_u64": lambda n: [workloads.fib_inputs(n, seed=11)],
    "fib_u32": lambda n: [workloads.fib_inputs(n, seed=12).astype(np.uint32)],
    "collatz_u64": lambda n: [workloads.collatz_inputs(n, seed=13)],
    "factorial_u64": lambda n: [workloads.uniform_int(14, n, 0, 25)],
    "jump4_u32": lambda n: [workloads.uniform_int(15, n, 0, 6).astype(np.uint32)],
    "int_mix_u64": lambda n: int_inputs("int_mix_u64", n, 16),
    "int_mix_u32": lambda n: int_inputs("int_mix_u32", n, 17),
    "fp_mix_f32": lambda n: [np.random.default_rng(18).uniform(-900, 900, n).astype(np.float32),
                             np.random.default_rng(19).uniform(-50, 50, n).astype(np.float32)],
    "fp_mix_f64": lambda n: [np.random.default_rng(20).uniform(-900, 900, n),
                             np.random.default_rng(21).uniform(-50, 50, n)],
    # path tracer: the first n samples of a 40x30 frame at 8 spp
    "smallpt9_f32": lambda n: [a[:n] for a in workloads.pt_inputs(40, 30, 8)],
    "threefry2x32_u32": lambda n: workloads.threefry_inputs(n, seed=22),
    "cubic_bisect_f32": lambda n: workloads.cubic_inputs(n, np.float32, seed=23),
    "cubic_bisect_f64": lambda n: workloads.cubic_inputs(n, np.float64, seed=24),
}


@pytest.mark.parametrize("mname,mode,params", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("name", list(CASES))
def test_parity_vs_oracle(gpu, name, mname, mode, params):
    for n in (1, 63, 65, 4099):
        ins = CASES[name](n)
        want, err = O.run(name, ins)
        assert err == 0
        got = gpu_run(name, mode, params, ins).outputs
        for g, w in zip(got, want):
            assert same(g, w), (name, mname, n)


def test_empty_input(gpu):
    r = loom_amd.run("fib_u64", abi.MODE_COMPACT, [np.zeros(0, np.uint64)])
    assert len(r.outputs[0]) == 0
    r = loom_amd.run("fib_u64", abi.MODE_HYBRID, [np.zeros(0, np.uint64)])
    assert len(r.outputs[0]) == 0


FIB = [0, 1]
for _ in range(89):
    FIB.append(FIB[-1] + FIB[-2])
FIB = np.array(FIB, dtype=np.uint64)


@pytest.mark.parametrize("mode,params", [(abi.MODE_COMPACT, {}),
                                         (abi.MODE_HYBRID, dict(threshold=32, regroup=True)),
                                         (abi.MODE_HYBRID, {})])
def test_fib_full_config(gpu, mode, params):
    """fib config: 2^20 threads, n ~ U[0,90], bit-exact vs the closed form."""
    n = workloads.fib_inputs(1 << 20)
    r = gpu_run("fib_u64", mode, params, [n])
    assert np.array_equal(r.outputs[0], FIB[n.astype(np.int64)])
    assert 0 < r.active_lane_fraction <= 1


def collatz_ref(x):
    c = 0
    while x > 1:
        x = x // 2 if x % 2 == 0 else 3 * x + 1
        c += 1
    return c


def test_collatz_full_config_forms_agree(gpu):
    """Collatz config: 2^24 threads; compact and hybrid agree bit for bit,
    a seeded sample matches brute force, and the first 2^16 match the oracle."""
    x = workloads.collatz_inputs(1 << 24)
    a = gpu_run("collatz_u64", abi.MODE_COMPACT, {}, [x]).outputs[0]
    b = gpu_run("collatz_u64", abi.MODE_HYBRID, dict(threshold=32, regroup=True), [x]).outputs[0]
    assert np.array_equal(a, b)
    idx = workloads.uniform_int(99, 400, 0, (1 << 24) - 1).astype(np.int64)
    assert [int(a[i]) for i in idx] == [collatz_ref(int(x[i])) for i in idx]
    (w,), _ = O.run("collatz_u64", [x[:1 << 16]])
    assert np.array_equal(a[:1 << 16], w)


@pytest.mark.parametrize("mname,mode,params", MODES, ids=[m[0] for m in MODES])
def test_smallpt9_frame_vs_oracle(gpu, mname, mode, params):
    """A whole 64x48x16spp frame: per-sample radiance bit-exact vs the oracle,
    image finite and non-negative."""
    w, h, spp = 64, 48, 16
    ins = workloads.pt_inputs(w, h, spp)
    want, err = O.run("smallpt9_f32", ins)
    assert err == 0
    got = gpu_run("smallpt9_f32", mode, params, ins).outputs
    for g, o in zip(got, want):
        assert same(g, o), mname
    img = workloads.pt_image(*got, w, h, spp)
    assert np.isfinite(img).all() and (img >= 0).all()


@pytest.mark.parametrize("name,n", [("collatz_u64", 1 << 18), ("smallpt9_f32", 40 * 30 * 8),
                                    ("fib_u64", 100000)])
@pytest.mark.parametrize("live", [0, 48])
def test_regroup_conserves_threads(gpu, name, n, live):
    """Queue regrouping moves threads, never loses or duplicates them: every
    spilled row is filled again, every input thread is admitted once, and
    the outputs equal the in-place run's."""
    ins = CASES[name](n)
    a = gpu_run(name, abi.MODE_HYBRID, dict(live_threshold=live, regroup=True,
                                            waves_per_cu=2, queue_capacity=1 << 16), ins)
    b = gpu_run(name, abi.MODE_HYBRID, dict(live_threshold=live, regroup=False), ins)
    for x, y in zip(a.outputs, b.outputs):
        assert same(x, y)
    c = a.counters
    assert c[abi.C_SPILL_ROWS] == c[abi.C_FILL_ROWS]
    assert c[abi.C_ADMITTED] == n
    assert c[abi.C_SPILL_FAILED] == 0
    assert c[abi.C_THREAD_STEPS] == b.counters[abi.C_THREAD_STEPS]
