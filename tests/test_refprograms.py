"""The reference's own test programs (src/tests/programs.hpp), compiled
unmodified from the reference checkout against this repository's front-end
(oracle/build_ref.py): every instantiation its VM test suites use lowers, the
oracle interpreter running them agrees with the reference's own validators
(fn::val_* and the std functors) on the inputs its generators draw, and the
invariant programs hold.  Needs no GPU; test_gpu_refprograms.py runs the same
programs on the GPU."""
import ctypes as C

import numpy as np
import pytest

import loom_amd
from tests import oracle_lib as O
from tests import refprograms as P
from tests.test_refkernels import REF_LIB, load_ref_kernels

pytestmark = pytest.mark.skipif(not (REF_LIB.exists() and O.REF_PATH.exists()),
                                reason="reference kernels not built (no reference checkout)")

N = 4096


def test_every_program_registered():
    load_ref_kernels()
    names = {loom_amd.kernel(n).name for n in P.NAMES}
    assert len(names) == len(P.NAMES) == 80
    assert loom_amd.kernel(P.RNG).outputs == [np.uint32] * 4 + [np.float32] * 2
    for name in P.NAMES:
        k = loom_amd.kernel(name)
        _, ins, outs = O.io(name, O.REF_PATH)
        assert k.inputs == ins and k.outputs == outs, name


@pytest.mark.parametrize("name", P.NAMES)
def test_oracle_matches_reference_validator(name):
    cols = P.inputs(name, N)
    (got, *rest), err = O.run(name, cols, path=O.REF_PATH)
    assert err == 0 and not rest
    want = P.validate(name, cols)
    if want is None:
        # invariant programs: every output non-zero (src/common/test.hpp:636-648)
        assert np.count_nonzero(got) == N, name
    elif P.PROGRAMS[P.split(name)[0]][1] == "fuzzy":
        # at most 1% beyond 100 ULP (src/tests/vm_test_suite_utils.hpp:172-176)
        assert P.ulp_mismatches(got, want) <= 0.01 * N, name
    else:
        assert np.array_equal(got, want), name


def test_validators_pin_known_values():
    """The validators against values fixed by arithmetic (so a wrong
    validator entry cannot pass the comparison above by agreeing with a
    wrong interpreter)."""
    fib = np.array([0, 1, 2, 10, 50, 80, 93], np.uint64)
    assert P.validate("ref_fib_test_u64", [fib]).tolist() == [
        0, 1, 1, 55, 12586269025, 23416728348467685, 12200160415121876738]
    fac = np.array([0, 1, 5, 12], np.uint32)
    assert P.validate("ref_fac_test_u32", [fac]).tolist() == [1, 1, 120, 479001600]
    j = np.arange(5, dtype=np.uint8)
    assert P.validate("ref_jmp4_test_u8", [j]).tolist() == [1, 3, 7, 11, 4]
    a = np.array([2, 10], np.uint16)
    assert P.validate("ref_lshift_test_u16", [a, a]).tolist() == [8, 10 << 10]
    assert P.validate("ref_bit_not_test_u8", [np.array([0, 0xF0], np.uint8)]).tolist() == [
        0xFF, 0x0F]
    assert P.validate("ref_increment_i8", [np.array([-128, 126], np.int8)]).tolist() == [
        -127, 127]


def test_wrong_validator_name_is_rejected():
    L = O.lib(O.REF_PATH)
    L.ref_validate.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.c_void_p, C.c_uint64]
    x = np.zeros(4, np.uint32)
    ip = (C.c_void_p * 16)(x.ctypes.data)
    assert L.ref_validate(b"ref_no_such_program", ip, x.ctypes.data, 4) == -1


def test_rng_macros_match_reference_functions():
    """The reference's Threefry / bump_ctr / gen_canonical IR macros through
    the oracle equal its C++ functions bit for bit (the check of
    src/tests/vm_smallpt.cpp:358-376), including counters that overflow."""
    cols = P.rng_inputs(N)
    got, err = O.run(P.RNG, cols, path=O.REF_PATH)
    assert err == 0
    want = P.rng_reference(*cols)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))


def test_threefry_kernel_matches_reference_function():
    """This repository's own Threefry-2x32 kernel (threefry2x32_u32) against
    the reference's C++ Threefry-2x32 with the same key: the block words
    agree bit for bit.  Its counter output is a 64-bit add (c1:c0 +
    39916801), which equals the reference's bump_ctr except where the low
    word wraps (bump_ctr then adds the wrapped low word to the high word and
    zeroes the low word, src/common/random.hpp:93-103)."""
    c0, c1 = P.rng_inputs(N)
    (x0, x1, n0, n1), err = O.run("threefry2x32_u32", [c0, c1])
    assert err == 0
    w0, w1, b0, b1, _, _ = P.rng_reference(c0, c1)
    assert np.array_equal(x0, w0) and np.array_equal(x1, w1)
    wraps = c0.astype(np.uint64) + 39916801 > 0xFFFFFFFF
    assert wraps.any() and (~wraps).any()
    assert np.array_equal(n0[~wraps], b0[~wraps]) and np.array_equal(n1[~wraps], b1[~wraps])
    assert np.array_equal(b0[wraps], np.zeros(wraps.sum(), np.uint32))


@pytest.mark.parametrize("name", P.JMP4)
def test_jmp4_past_the_table_matches_reference_validator(name):
    """The reference's validator returns the index itself for indexes past
    the table; the op contract (such a jump halts) gives the same outputs."""
    cols = P.jmp4_wide_inputs(name)
    (got,), err = O.run(name, cols, path=O.REF_PATH)
    assert err == 0
    assert np.array_equal(got, P.validate(name, cols))


def test_reference_fib_at_full_size():
    """The fib workload at its full size (2^20 threads, n in [0, 90]) through
    the oracle running the reference's fib_test equals the reference's
    val_fib, and this repository's fib_u64 kernel agrees with both."""
    from loom_amd import workloads
    n = workloads.fib_inputs()
    want = P.validate("ref_fib_test_u64", [n])
    (got,), err = O.run("ref_fib_test_u64", [n], path=O.REF_PATH)
    assert err == 0 and np.array_equal(got, want)
    (own,), err = O.run("fib_u64", [n])
    assert err == 0 and np.array_equal(own, want)
