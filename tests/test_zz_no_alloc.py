"""loom_run_noalloc, the run entry point that allocates nothing (its hybrid
control block and regroup queues live in the library's static device
scratch array), through the C ABI against loom_run on the same inputs: bit
for bit, on two calls in a row, the second with a smaller input, in the
compact form and in the hybrid form with and without queue regrouping.
Collected last: it also runs after every other GPU test of the process."""
import ctypes as C

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def run_both(name, mode, ins, params):
    k = loom_amd.kernel(name)
    n = len(ins[0])
    din = [loom_amd.DeviceArray.from_numpy(np.ascontiguousarray(a, dtype=dt))
           for a, dt in zip(ins, k.inputs)]
    ref = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
    got = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
    loom_amd.run(k, mode, din, ref, params=params)
    ip = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in din])
    op = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in got])
    st = abi.RunStats()
    abi.check(abi.load().loom_run_noalloc(k.index, mode, ip, op, n, C.byref(params), C.byref(st)),
              "loom_run_noalloc")
    assert st.errors == 0
    for r, g in zip(ref, got):
        assert np.array_equal(bits(r.numpy()), bits(g.numpy())), (name, mode, n)
    return st


CASES = [
    ("collatz_u64", abi.MODE_HYBRID, dict(live_threshold=48)),
    ("collatz_u64", abi.MODE_HYBRID, dict(live_threshold=32, regroup=True, refill=False)),
    ("fib_u64", abi.MODE_HYBRID, dict(live_threshold=32, regroup=True)),
    ("smallpt9_f32", abi.MODE_HYBRID, dict(live_threshold=48)),
    ("fib_u64", abi.MODE_COMPACT, dict()),
]


@pytest.mark.parametrize("name,mode,kw", CASES)
def test_noalloc_equals_run_twice(gpu, name, mode, kw):
    params = loom_amd.make_params(mode=mode, **kw)
    sizes = (50000, 3001) if name != "smallpt9_f32" else ((32, 24, 8), (16, 8, 4))
    for size in sizes:
        if name == "smallpt9_f32":
            ins = workloads.pt_inputs(*size)
        elif name == "fib_u64":
            ins = [workloads.fib_inputs(size, seed=70)]
        else:
            ins = [workloads.collatz_inputs(size, seed=71)]
        print(f"[gpu] noalloc {name} mode={mode} {kw} n={len(ins[0])}", flush=True)
        run_both(name, mode, ins, params)


def test_noalloc_rejects_queues_beyond_scratch(gpu):
    """A regrouping run whose queues exceed the scratch array fails with
    LOOM_E_CAPACITY before launching anything."""
    k = loom_amd.kernel("smallpt9_f32")
    n = 1 << 22
    params = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=True, refill=False,
                                  queue_capacity=1 << 22)
    din = [loom_amd.DeviceArray(n, dt) for dt in k.inputs]
    dout = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
    ip = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in din])
    op = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in dout])
    rc = abi.load().loom_run_noalloc(k.index, abi.MODE_HYBRID, ip, op, n, C.byref(params), None)
    assert rc == abi.E_CAPACITY
