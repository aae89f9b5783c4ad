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
