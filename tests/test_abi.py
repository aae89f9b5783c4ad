"""The C ABI library loads, exports every declared entry point, and reports
kernel structure that matches hand derivations from the reference rules
(core.hpp:686-1031).  No GPU call is made here."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import loom_amd
from loom_amd import abi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "loom_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(loom_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = abi.load()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding covers them all
    assert set(names) <= set(abi.SIGNATURES)


def test_fib_structure():
    k = loom_amd.kernel("fib_u64")
    # blocks: START (le, cmov, pjmp), work (3 mov + fall-through jmp),
    # loop (add, mov, mov, add, le + pjmp)
    assert k.n_blocks == 3
    assert k.block_ops == [3, 4, 6]
    # registers n, c, cmp, p, i, t (u64) + the u64 PC: 56 bytes (core.hpp:1027)
    assert k.n_registers == 6 and k.state_size == 56
    # live-in at START {n}, at work {n}, at loop {p, c, i, n}; +4 B thread index
    assert k.block_live_bytes == [12, 12, 36]
    assert k.inputs == [np.uint64] and k.outputs == [np.uint64]


def test_other_structures():
    c = loom_amd.kernel("collatz_u64")
    assert c.n_blocks == 6            # START step even odd3 next done
    assert c.outputs == [np.uint32]
    j = loom_amd.kernel("jump4_u32")
    assert j.n_blocks == 5
    assert j.block_ops == [1, 2, 2, 2, 2]
    m = loom_amd.kernel("int_mix_u64")
    assert m.n_blocks == 1 and m.block_ops == [41]   # straight-line
    assert len(m.outputs) == 6


def test_kernel_lookup_errors():
    with pytest.raises(KeyError):
        loom_amd.kernel("no_such_kernel")
    assert abi.load().loom_kernel_find(b"no_such_kernel") == -1


def test_calls_before_init_fail_cleanly():
    # without a successful loom_init every device entry point reports an error
    import ctypes as C
    lib = abi.load()
    if loom_amd._initialised:
        pytest.skip("device already initialised in this process")
    p = C.c_void_p()
    assert lib.loom_malloc(C.byref(p), 16) == 1002
    assert b"loom_init" in lib.loom_last_error()


def test_kernel_listing():
    """The lowered listing of fib names its three blocks, their successors
    and the loop's instructions (no device needed)."""
    from loom_amd import inspector as inspect
    text = inspect.listing("fib_u64")
    assert "3 blocks" in text
    assert "block 0 START" in text and "block 1 work" in text and "block 2 loop" in text
    assert "successors: work HALT" in text and "successors: HALT loop" in text
    assert "= add.u64 p c" in text and "pjump cmp -> HALT loop" in text
    lib = abi.load()
    need = C.c_size_t()
    assert lib.loom_kernel_dump(999, None, 0, C.byref(need)) == abi.E_KERNEL
    small = C.create_string_buffer(8)
    assert lib.loom_kernel_dump(0, small, 8, C.byref(need)) == 0
    assert len(small.value) == 7 and need.value > 8
