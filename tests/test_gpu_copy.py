"""Background copies on the second stream: tickets complete in order, data
round-trips through device memory, waits return, bad tickets are errors."""
import ctypes as C

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, serve

pytestmark = pytest.mark.gpu


def test_copy_round_trip_and_tickets(gpu):
    lib = abi.load()
    n = 1 << 20
    src = serve.Pinned(n, np.uint32)
    src.array[:] = np.arange(n, dtype=np.uint32) * np.uint32(2654435761)
    dst = serve.Pinned(n, np.uint32)
    dev = loom_amd.DeviceArray(n, np.uint32)
    t1, t2 = C.c_uint64(), C.c_uint64()
    print("[gpu] copy tickets", flush=True)
    abi.check(lib.loom_copy_async(dev.ptr, src.ptr, 4 * n, abi.COPY_H2D, C.byref(t1)), "h2d")
    abi.check(lib.loom_copy_async(dst.ptr, dev.ptr, 4 * n, abi.COPY_D2H, C.byref(t2)), "d2h")
    assert t2.value == t1.value + 1
    abi.check(lib.loom_copy_wait(t2), "wait")
    done = C.c_int()
    abi.check(lib.loom_copy_status(t1, C.byref(done)), "status")
    assert done.value == 1  # an earlier ticket is complete once a later one is
    assert np.array_equal(dst.array, src.array)
    assert lib.loom_copy_wait(t2.value + 100) == abi.E_ARG
    assert lib.loom_copy_async(dst.ptr, dev.ptr, 4, 7, C.byref(t1)) == abi.E_ARG
    # an empty copy still gets a ticket that completes
    t3 = C.c_uint64()
    abi.check(lib.loom_copy_async(None, None, 0, abi.COPY_H2D, C.byref(t3)), "empty")
    abi.check(lib.loom_copy_wait(t3), "wait empty")
