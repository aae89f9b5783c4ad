"""Regrouping inside a workgroup (loom_params.regroup = LOOM_REGROUP_WORKGROUP):
the 1024 (or, with 8-wave workgroups, 512) threads of a workgroup are sorted
by block through LDS between run phases, their live registers moving with
them.  Every output equals the oracle's bit for bit, and the useful work the
device counts (active operations and flops, thread-block steps) equals the
compact form's: the same blocks run on the same threads, only grouped
differently."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


CASES = {
    "fib_u64": lambda: [workloads.fib_inputs(70000, seed=81)],
    "collatz_u64": lambda: [workloads.collatz_inputs(60000, seed=82)],
    "smallpt9_f32": lambda: workloads.pt_inputs(48, 32, 16),
    "cubic_bisect_f32": lambda: workloads.cubic_inputs(20000),
    "fp_mix_f64": lambda: [np.random.default_rng(83).uniform(-900, 900, 5000),
                           np.random.default_rng(84).uniform(-900, 900, 5000)],
    "jump4_u32": lambda: [np.random.default_rng(85).integers(0, 9, 3000, dtype=np.uint32)],
}
FORMS = [("successor-major thr16 phase64", dict(reconv=abi.RECONV_SUCCESSOR_MAJOR, threshold=16,
                                                regroup_phase=64)),
         ("successor-major thr0 phase8", dict(reconv=abi.RECONV_SUCCESSOR_MAJOR, threshold=0,
                                              regroup_phase=8)),
         ("sweep phase32", dict(reconv=abi.RECONV_SWEEP, regroup_phase=32)),
         ("8 waves successor-major thr8 phase16", dict(reconv=abi.RECONV_SUCCESSOR_MAJOR,
                                                       threshold=8, regroup_phase=16,
                                                       regroup_waves=8)),
         ("8 waves sweep phase64", dict(reconv=abi.RECONV_SWEEP, regroup_phase=64,
                                        regroup_waves=8))]
WORK = (abi.C_ACTIVE_OPS, abi.C_THREAD_STEPS, abi.C_ACTIVE_FLOPS)


@pytest.mark.parametrize("name", list(CASES))
def test_workgroup_regroup_matches_oracle(gpu, name):
    ins = CASES[name]()
    k = loom_amd.kernel(name)
    want, err = O.run(name, ins)
    assert err == 0
    base = loom_amd.run(k, abi.MODE_COMPACT, ins, params=loom_amd.make_params(
        mode=abi.MODE_COMPACT, reconv=abi.RECONV_SWEEP))
    for fname, params in FORMS:
        print(f"[gpu] workgroup regroup {name} {fname} n={len(ins[0])}", flush=True)
        p = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=abi.REGROUP_WORKGROUP, **params)
        r = loom_amd.run(k, abi.MODE_HYBRID, ins, params=p)
        assert r.errors == 0, (name, fname)
        for g, w in zip(r.outputs, want):
            assert np.array_equal(bits(g), bits(w)), (name, fname)
        for c in WORK:
            assert r.counters[c] == base.counters[c], (name, fname, c)


def test_workgroup_regroup_small_and_empty(gpu):
    """Fewer threads than one workgroup's lanes."""
    k = loom_amd.kernel("collatz_u64")
    for waves in (16, 8):
        for n in (1, 63, 511, 1000):
            x = [workloads.collatz_inputs(n, seed=86 + n)]
            want, _ = O.run("collatz_u64", x)
            p = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=abi.REGROUP_WORKGROUP,
                                     reconv=abi.RECONV_SWEEP, regroup_waves=waves)
            r = loom_amd.run(k, abi.MODE_HYBRID, x, params=p)
            assert np.array_equal(r.outputs[0], want[0]), (waves, n)
            assert r.grid_waves % waves == 0, (waves, n)


def test_workgroup_regroup_rejected_with_topo(gpu):
    k = loom_amd.kernel("fib_u64")
    p = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=abi.REGROUP_WORKGROUP,
                             reconv=abi.RECONV_TOPO)
    with pytest.raises(abi.LoomError):
        loom_amd.run(k, abi.MODE_HYBRID, [workloads.fib_inputs(100)], params=p)


def test_workgroup_regroup_rejects_other_sizes(gpu):
    k = loom_amd.kernel("fib_u64")
    for waves in (4, 12, 32):
        p = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=abi.REGROUP_WORKGROUP,
                                 reconv=abi.RECONV_SWEEP, regroup_waves=waves)
        with pytest.raises(abi.LoomError):
            loom_amd.run(k, abi.MODE_HYBRID, [workloads.fib_inputs(100)], params=p)
