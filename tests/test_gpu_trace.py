"""Step traces on the GPU (loom_kernel_trace) against the oracle's, row by
row: the block run, the error bits and every register's bits after every
block, for every built-in kernel on seeded inputs; the trace's own
consistency (next block = following row's block, HALT last, last row =
the outputs of a whole run); argument errors."""
import ctypes as C

import numpy as np
import pytest

import loom_amd
from loom_amd import abi
from loom_amd import inspector as I
from tests import oracle_lib as O
from tests.test_gpu_parity import CASES

pytestmark = pytest.mark.gpu

# threads and step limit per kernel: long enough that most threads halt,
# short enough that truncated threads are also compared
SIZES = {"smallpt9_f32": (24, 3000), "collatz_u64": (64, 600)}


@pytest.mark.parametrize("name", list(CASES))
def test_trace_matches_oracle(gpu, name):
    n, max_steps = SIZES.get(name, (65, 2000))
    ins = CASES[name](n)
    print(f"[gpu] trace {name} n={n} max_steps={max_steps}", flush=True)
    tr = I.trace(name, ins, max_steps)
    names = [g["name"] for g in tr.registers]
    halted = 0
    for t in range(n):
        want = O.trace(name, ins, t, max_steps)
        steps = int(tr.steps[t])
        assert steps == len(want), (name, t)
        for s, (labels, err, regs) in enumerate(want):
            assert tr.labels[int(tr.block[t, s])] in labels, (name, t, s)
            assert int(tr.errors[t, s]) == err, (name, t, s)
            got = {nm: int(b) for nm, b in zip(names, tr.bits[t, s])}
            assert got == regs, (name, t, s)
        # the trace's own chain: each row's next block is the next row's block
        assert np.array_equal(tr.next_block[t, :steps - 1], tr.block[t, 1:steps])
        halted += int(tr.next_block[t, steps - 1]) == abi.TRACE_HALT
    assert halted > 0


def test_last_rows_are_run_outputs(gpu):
    name = "cubic_bisect_f32"
    ins = CASES[name](200)
    tr = I.trace(name, ins, 4000)
    outs = loom_amd.run(name, abi.MODE_COMPACT, ins).outputs
    regs = [r for r, g in enumerate(tr.registers) if g["output"]]
    assert len(regs) == len(outs)
    for t in range(200):
        s = int(tr.steps[t]) - 1
        assert int(tr.next_block[t, s]) == abi.TRACE_HALT
        for r, col in zip(regs, outs):
            assert tr.value(t, s, r).tobytes() == col[t:t + 1].tobytes()


def test_fib_trace_table(gpu):
    tr = I.trace("fib_u64", [np.array([6], np.uint64)], 50)
    text = I.format_trace(tr, 0, ["c"])
    rows = [ln.split("|") for ln in text.splitlines()[2:-1]]
    assert [r[0].strip() for r in rows] == ["START", "work"] + ["loop"] * 5
    assert [int(r[1]) for r in rows[2:]] == [1, 2, 3, 5, 8]
    assert text.splitlines()[-1] == "HALT"


def test_trace_argument_errors(gpu):
    lib = abi.load()
    k = loom_amd.kernel("fib_u64")
    col = loom_amd.DeviceArray.from_numpy(np.arange(4, dtype=np.uint64))
    rows = loom_amd.DeviceArray(4 * 8 * abi.trace_words(k.n_registers), np.uint64)
    steps = loom_amd.DeviceArray(4, np.uint32)
    ins = (C.c_void_p * abi.MAX_IO)(col.ptr)
    assert lib.loom_kernel_trace(k.index, ins, 4, 0, rows.ptr, rows.nbytes, steps.ptr) == abi.E_ARG
    assert lib.loom_kernel_trace(k.index, ins, 4, 1 << 17, rows.ptr, rows.nbytes,
                                 steps.ptr) == abi.E_ARG
    assert lib.loom_kernel_trace(k.index, ins, 4, 9, rows.ptr, rows.nbytes,
                                 steps.ptr) == abi.E_CAPACITY
    assert lib.loom_kernel_trace(k.index, ins, 1 << 20, 1, rows.ptr, rows.nbytes,
                                 steps.ptr) == abi.E_CAPACITY
    assert lib.loom_kernel_trace(k.index, None, 4, 8, rows.ptr, rows.nbytes,
                                 steps.ptr) == abi.E_ARG
    assert lib.loom_kernel_trace(k.index, ins, 0, 8, None, 0, None) == 0
    assert lib.loom_kernel_trace(k.index, ins, 4, 8, rows.ptr, rows.nbytes, steps.ptr) == 0
    assert list(steps.numpy()) == [1, 1, 3, 4]
