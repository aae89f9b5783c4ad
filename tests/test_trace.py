"""Step traces without a GPU: the oracle's per-block trace against closed
forms, the register and label metadata of the product library against the
oracle's, and the table format of a trace."""
import numpy as np
import pytest

import loom_amd
from loom_amd import inspector as I
from tests import oracle_lib as O

FIB = [0, 1]
for _ in range(92):
    FIB.append(FIB[-1] + FIB[-2])

ALL = [k.name for k in loom_amd.kernels()]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 10, 40, 93])
def test_oracle_trace_fib_closed_form(n):
    """fib(n): START, then for n > 1 `work` and n-1 iterations of `loop`; after
    iteration k the loop holds p = F(k), c = F(k+1), i = k+2, t = c."""
    tr = O.trace("fib_u64", [np.array([n], np.uint64)], 0, 1000)
    labels = [sorted(l) for l, _, _ in tr]
    if n <= 1:
        assert labels == [["START"]]
        assert tr[0][2]["c"] == n and tr[0][2]["cmp"] == 1
        return
    assert labels == [["START"], ["work"]] + [["loop"]] * (n - 1)
    assert tr[1][2] == {"n": n, "cmp": 0, "c": 1, "p": 0, "i": 2, "t": 0}
    for k in range(1, n):
        r = tr[1 + k][2]
        assert (r["p"], r["c"], r["i"], r["t"]) == (FIB[k], FIB[k + 1], k + 2, FIB[k + 1])
        assert r["cmp"] == (1 if k + 2 <= n else 0)
    assert all(e == 0 for _, e, _ in tr)


def test_oracle_trace_stops_at_max_steps():
    tr = O.trace("fib_u64", [np.array([50], np.uint64)], 0, 7)
    assert len(tr) == 7 and tr[-1][2]["i"] == 7


@pytest.mark.parametrize("name", ["collatz_u64", "cubic_bisect_f32", "threefry2x32_u32"])
def test_oracle_trace_ends_in_run_outputs(name):
    """The last row of each thread holds what a whole run outputs."""
    from tests.test_gpu_parity import CASES
    ins = CASES[name](5)
    outs, err = O.run(name, ins, threads=1)
    assert err == 0
    out_regs = [g["name"] for g in I.registers(name) if g["output"]]
    for t in range(5):
        last = O.trace(name, ins, t, 100000)[-1][2]
        for reg, col in zip(out_regs, outs):
            assert last[reg] == int(col[t:t + 1].view(np.uint32 if col.itemsize == 4
                                                       else np.uint64)[0])


def test_registers_and_labels_of_fib():
    regs = I.registers("fib_u64")
    assert [g["name"] for g in regs] == ["n", "c", "cmp", "p", "i", "t"]
    assert all(g["dtype"] == np.uint64 and g["width"] == 8 and g["ntypes"] == 1 for g in regs)
    assert [g["name"] for g in regs if g["input"]] == ["n"]
    assert [g["name"] for g in regs if g["output"]] == ["c"]
    assert not next(g for g in regs if g["name"] == "n")["written"]
    assert I.block_labels("fib_u64") == ["START", "work", "loop"]


@pytest.mark.parametrize("name", ALL)
def test_register_names_match_oracle(name):
    """The product's registers are exactly the oracle's (by name), so a GPU
    trace compares with the oracle's register by register."""
    tr = O.trace(name, [np.zeros(1, dt) for dt in loom_amd.kernel(name).inputs], 0, 1)
    assert {g["name"] for g in I.registers(name)} == set(tr[0][2])
    assert len(I.block_labels(name)) == loom_amd.kernel(name).n_blocks


def test_register_query_errors():
    from loom_amd import abi
    lib = abi.load()
    k = loom_amd.kernel("fib_u64").index
    assert lib.loom_kernel_register(k, 6, None, 0, None, None, None, None) == abi.E_ARG
    assert lib.loom_kernel_register(k, -1, None, 0, None, None, None, None) == abi.E_ARG
    assert lib.loom_kernel_register(999, 0, None, 0, None, None, None, None) == abi.E_KERNEL
    assert lib.loom_kernel_register(k, 0, None, 0, None, None, None, None) == 0


def test_format_trace_table():
    regs = I.registers("fib_u64")
    n, steps = 1, 4
    bits = np.zeros((n, steps, len(regs)), np.uint64)
    tr = O.trace("fib_u64", [np.array([3], np.uint64)], 0, steps)
    for s, (_, _, r) in enumerate(tr):
        bits[0, s] = [r[g["name"]] for g in regs]
    blocks = np.array([[0, 1, 2, 2]], np.uint32)
    nxt = np.array([[1, 2, 2, 0xFFFF]], np.uint32)
    t = I.Trace(kernel="fib_u64", registers=regs, labels=I.block_labels("fib_u64"),
                steps=np.array([len(tr)], np.uint32), block=blocks, next_block=nxt,
                errors=np.zeros((n, steps), np.uint64), bits=bits)
    text = I.format_trace(t, 0, ["p", "c"], width=8).splitlines()
    assert text[0].split("|")[:3] == ["BLOCK ", "p      ", "c      "]
    assert [ln.split("|")[0].strip() for ln in text[2:6]] == ["START", "work", "loop", "loop"]
    assert [ln.split("|")[2].strip() for ln in text[4:6]] == ["1", "2"]
    assert text[-1] == "HALT"
