"""The reference's own path-tracing kernels (src/smallpt/vm/pt.hpp:
pt_const, pt2_const), included from the reference checkout and lowered by
this repository's front-end (oracle/build_ref.py builds them into
oracle/_ref/ where the checkout exists): the drop-in check that existing
kernels run unchanged.  This file needs no GPU; test_gpu_refkernels.py runs
them on the GPU against the oracle."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import loom_amd
from loom_amd import abi
from tests import oracle_lib as O

ROOT = Path(__file__).resolve().parent.parent
REF_LIB = ROOT / "oracle" / "_ref" / "libloom_refkernels.so"
# block counts implied by the reference's stream sizes 288 = 16*ceil(0.4*B)
# and 352 = 16*ceil(0.4*B) (src/benchmarks/pt.cpp:62-68; bm-results.md:40,82)
KERNELS = {"ref_pt_const_b1_t0.0001_d2": (43, 44, 45), "ref_pt2_const_b5_t0.0005_d20": (53, 54, 55)}

pytestmark = pytest.mark.skipif(not (REF_LIB.exists() and O.REF_PATH.exists()),
                                reason="reference kernels not built (no reference checkout)")

_handle = None


def load_ref_kernels():
    """Register the reference kernels with the product library."""
    global _handle
    abi.load()
    if _handle is None:
        _handle = C.CDLL(str(REF_LIB))
    return _handle


def frame_inputs(w=40, h=40):
    """ppp = x << 16 | y over the frame (reference smallpt/vm/enc.hpp:30-34)."""
    return np.array([(x << 16) | y for x in range(w) for y in range(h)], np.uint32)


@pytest.mark.parametrize("name", list(KERNELS))
def test_reference_kernel_lowers(name):
    load_ref_kernels()
    k = loom_amd.kernel(name)
    assert k.inputs == [np.uint32]
    assert k.outputs == [np.float32, np.float32, np.float32, np.uint32, np.uint32]
    assert k.n_blocks in KERNELS[name]
    _, ins, outs = O.io(name, O.REF_PATH)
    assert ins == k.inputs and outs == k.outputs


def test_reference_kernel_oracle_row():
    """A few pixels through the oracle: finite radiance, the pixel index
    passed through, at least the minimum 64 samples' worth of rays."""
    ppp = frame_inputs()[::97]
    (l0, l1, l2, rc, out_ppp), err = O.run("ref_pt_const_b1_t0.0001_d2", [ppp], path=O.REF_PATH)
    assert err == 0
    assert np.array_equal(out_ppp, ppp)
    assert np.isfinite(l0).all() and np.isfinite(l1).all() and np.isfinite(l2).all()
    assert (rc >= 64).all()
