"""The compact form with its register file in device memory
(loom_params.register_file = LOOM_REGFILE_MEMORY): one column per register,
one row per thread, each block loading what it reads and storing what stays
live.  Every output equals the oracle's (and the VGPR form's) bit for bit,
and the device's work counters equal the VGPR form's: the same blocks run on
the same lanes."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


CASES = {
    "fib_u64": lambda: [workloads.fib_inputs(50000, seed=61)],
    "collatz_u64": lambda: [workloads.collatz_inputs(40000, seed=62)],
    "smallpt9_f32": lambda: workloads.pt_inputs(32, 24, 8),
    "cubic_bisect_f64": lambda: workloads.cubic_inputs(8000, np.float64),
    "fp_mix_f32": lambda: [np.random.default_rng(63).uniform(-900, 900, 9000).astype(np.float32),
                           np.random.default_rng(64).uniform(-900, 900, 9000).astype(np.float32)],
    "roots_f32": lambda: [np.random.default_rng(65).uniform(0, 1e6, 7000).astype(np.float32)],
}


@pytest.mark.parametrize("name", list(CASES))
def test_memory_register_file_matches_oracle(gpu, name):
    ins = CASES[name]()
    k = loom_amd.kernel(name)
    print(f"[gpu] memory register file {name} n={len(ins[0])}", flush=True)
    pm = loom_amd.make_params(mode=abi.MODE_COMPACT, reconv=abi.RECONV_SWEEP,
                              register_file=abi.REGFILE_MEMORY)
    rm = loom_amd.run(k, abi.MODE_COMPACT, ins, params=pm)
    pv = loom_amd.make_params(mode=abi.MODE_COMPACT, reconv=abi.RECONV_SWEEP)
    rv = loom_amd.run(k, abi.MODE_COMPACT, ins, params=pv)
    want, err = O.run(name, ins)
    assert err == 0 and rm.errors == 0
    for g, v, w in zip(rm.outputs, rv.outputs, want):
        assert np.array_equal(bits(g), bits(w)), name
        assert np.array_equal(bits(g), bits(v)), name
    assert rm.counters == rv.counters, name


def test_memory_register_file_rejected_elsewhere(gpu):
    """The memory register file is a compact-form option under the sweep
    policy; other combinations are argument errors, not silent fallbacks."""
    k = loom_amd.kernel("fib_u64")
    x = [workloads.fib_inputs(256, seed=66)]
    for mode, reconv in ((abi.MODE_COMPACT, abi.RECONV_SUCCESSOR_MAJOR),
                         (abi.MODE_HYBRID, abi.RECONV_SWEEP)):
        p = loom_amd.make_params(mode=mode, reconv=reconv, register_file=abi.REGFILE_MEMORY)
        with pytest.raises(abi.LoomError):
            loom_amd.run(k, mode, x, params=p)


def test_memory_register_file_resources(gpu):
    from loom_amd import inspector
    r = inspector.resources("smallpt9_f32")["compact-memory/sweep"]
    assert 0 < r["vgprs"] <= 128 and r["scratch_bytes"] <= 16, r
