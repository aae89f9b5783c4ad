"""Failure handling through the C ABI on the GPU.  Every condition is forced
from the host (a tiny dispatch budget, invalid arguments), so every kernel
still exits normally: the budget is checked by each wave after each block,
and a wave over budget leaves its loop."""
import ctypes as C

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,params", [(abi.MODE_COMPACT, {}),
                                         (abi.MODE_COMPACT, dict(reconv=abi.RECONV_TOPO)),
                                         (abi.MODE_HYBRID, {}),
                                         (abi.MODE_HYBRID, dict(regroup=True,
                                                                reconv=abi.RECONV_SUCCESSOR_MAJOR))])
def test_step_budget_raises(gpu, mode, params):
    """A dispatch budget far below what Collatz needs trips the step-budget
    flag; the run reports LOOM_E_RUNTIME and the next run is clean."""
    x = workloads.collatz_inputs(5000, seed=3)
    p = loom_amd.make_params(mode=mode, max_steps=50, **params)
    print(f"[gpu] step budget mode={mode} {params}", flush=True)
    with pytest.raises(abi.LoomError) as e:
        loom_amd.run("collatz_u64", mode, [x], params=p)
    assert e.value.code == abi.E_RUNTIME
    assert "0x2" in str(e.value)  # the step-budget bit
    r = loom_amd.run("collatz_u64", mode, [x], params=loom_amd.make_params(mode=mode, **params))
    assert r.errors == 0


def test_invalid_arguments(gpu):
    lib = abi.load()
    k = loom_amd.kernel("fib_u64")
    a = np.arange(10, dtype=np.uint64)
    out = np.zeros(10, np.uint64)
    ins = (C.c_void_p * abi.MAX_IO)(a.ctypes.data)
    outs = (C.c_void_p * abi.MAX_IO)(out.ctypes.data)
    bad = loom_amd.make_params(mode=abi.MODE_HYBRID, live_threshold=64)
    assert lib.loom_run_host(k.index, abi.MODE_HYBRID, ins, outs, 10, C.byref(bad), None) == abi.E_ARG
    bad = loom_amd.make_params(mode=abi.MODE_HYBRID, reconv=7)
    assert lib.loom_run_host(k.index, abi.MODE_HYBRID, ins, outs, 10, C.byref(bad), None) == abi.E_ARG
    assert lib.loom_run_host(k.index, 9, ins, outs, 10, None, None) == abi.E_ARG
    assert lib.loom_run_host(999, abi.MODE_HYBRID, ins, outs, 10, None, None) == abi.E_KERNEL
    assert lib.loom_run_host(k.index, abi.MODE_HYBRID, ins, outs, 1 << 33, None, None) == abi.E_CAPACITY
    nul = (C.c_void_p * abi.MAX_IO)()
    assert lib.loom_run(k.index, abi.MODE_HYBRID, nul, outs, 10, None, None) == abi.E_ARG
    # a NULL parameter block means the documented defaults
    assert lib.loom_run_host(k.index, abi.MODE_HYBRID, ins, outs, 10, None, None) == 0
    fib = [0, 1]
    for _ in range(8):
        fib.append(fib[-1] + fib[-2])
    assert list(out) == fib


def test_batch_rejects_foreign_kernel(gpu):
    b = loom_amd.Batch("int_mix")
    pt = loom_amd.kernel("smallpt9_f32")
    with pytest.raises(ValueError):
        b.add(pt, [], [])
    lib = abi.load()
    j = (abi.Job * 1)()
    j[0].kernel = pt.index
    j[0].begin, j[0].end = 0, 0
    assert lib.loom_batch_run(b.set, j, 1, None, None) == abi.E_KERNEL
    assert lib.loom_batch_run(b.set, j, 0, None, None) == abi.E_ARG
