"""The reference's path-tracing and vector macro kernels
(oracle/ref_macros.hpp) on the GPU, in the compact and hybrid forms, bit for
bit against the oracle (which tests/test_refmacros.py holds to the
reference's C++)."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi
from tests import oracle_lib as O
from tests import refmacros as M
from tests.test_refkernels import REF_LIB, load_ref_kernels

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (REF_LIB.exists() and O.REF_PATH.exists()),
                                 reason="reference kernels not built (no reference checkout)")]

N = 20000


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


@pytest.mark.parametrize("name", sorted(M.CHECKS))
@pytest.mark.parametrize("mode", [abi.MODE_COMPACT, abi.MODE_HYBRID])
def test_macro_kernel_gpu_equals_oracle(gpu, name, mode):
    load_ref_kernels()
    ins = M.inputs(name, N)
    print(f"[gpu] macro kernel {name} mode={mode}", flush=True)
    got = loom_amd.run(name, mode, ins).outputs
    want, err = O.run(name, ins, path=O.REF_PATH)
    for c, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(bits(g), bits(w)), (name, mode, c)
