"""Runs that collect no statistics launch, for a kernel with blocks below the
scalar-count threshold, kernels compiled without per-lane work counts.  Their
outputs must equal the oracle's bit for bit in every form that has them
(compact under each policy, hybrid, workgroup regrouping of both sizes), and
the counters of such a run stay zero."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

CASES = {
    "fib_u64": lambda: [workloads.fib_inputs(50000, seed=91)],
    "collatz_u64": lambda: [workloads.collatz_inputs(40000, seed=92)],
    "jump4_u32": lambda: [np.random.default_rng(93).integers(0, 9, 3000, dtype=np.uint32)],
    "int_mix_u32": lambda: [np.random.default_rng(94).integers(0, 2 ** 32, 5000, dtype=np.uint32),
                            np.random.default_rng(95).integers(0, 2 ** 32, 5000, dtype=np.uint32)],
}
FORMS = [("compact successor-major", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR)),
         ("compact topo", abi.MODE_COMPACT, dict(reconv=abi.RECONV_TOPO)),
         ("compact sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
         ("hybrid sweep", abi.MODE_HYBRID, dict(reconv=abi.RECONV_SWEEP)),
         ("hybrid successor-major", abi.MODE_HYBRID, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR,
                                                          threshold=8)),
         ("workgroup 16", abi.MODE_HYBRID, dict(regroup=abi.REGROUP_WORKGROUP,
                                                reconv=abi.RECONV_SWEEP, regroup_phase=32)),
         ("workgroup 8", abi.MODE_HYBRID, dict(regroup=abi.REGROUP_WORKGROUP,
                                               reconv=abi.RECONV_SWEEP, regroup_phase=64,
                                               regroup_waves=8))]


@pytest.mark.parametrize("name", list(CASES))
def test_uncounted_runs_match_oracle(gpu, name):
    ins = CASES[name]()
    k = loom_amd.kernel(name)
    want, err = O.run(name, ins)
    assert err == 0
    for fname, mode, params in FORMS:
        print(f"[gpu] uncounted {name} {fname}", flush=True)
        r = loom_amd.run(k, mode, ins, params=loom_amd.make_params(mode=mode, stats=False,
                                                                   **params))
        assert r.errors == 0, (name, fname)
        for g, w in zip(r.outputs, want):
            assert np.array_equal(g, w), (name, fname)
        assert r.counters[abi.C_ACTIVE_OPS] == 0 and r.counters[abi.C_THREAD_STEPS] == 0
