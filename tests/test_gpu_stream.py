"""The device-resident stream (loom_stream_*): registers stay on the GPU
between steps, finished threads' outputs are exchanged for new inputs on the
device.  Driven with the reference's stream protocol (StreamWrapper,
src/tests/stream_wrapper.hpp:33-80) and checked thread by thread against the
oracle (sequence numbers restore input order), with the bytes each exchange
moves proportional to the rows exchanged."""
import numpy as np
import pytest

from loom_amd import abi, workloads
from loom_amd.stream import Stream
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def cases():
    return {
        "fib_u64": [workloads.fib_inputs(9001, seed=21)],
        "collatz_u64": [workloads.collatz_inputs(12345, seed=22)],
        "smallpt9_f32": workloads.pt_inputs(24, 16, 8),
        "cubic_bisect_f32": workloads.cubic_inputs(3000),
    }


@pytest.mark.parametrize("name", ["fib_u64", "collatz_u64", "smallpt9_f32", "cubic_bisect_f32"])
@pytest.mark.parametrize("slots,live,reconv", [(256, 32, abi.RECONV_SWEEP),
                                               (1024, 0, abi.RECONV_SWEEP),
                                               (64, 48, abi.RECONV_SUCCESSOR_MAJOR),
                                               (512, 16, abi.RECONV_TOPO)])
def test_stream_matches_oracle(gpu, name, slots, live, reconv):
    ins = cases()[name]
    print(f"[gpu] stream {name} slots={slots} live={live} reconv={reconv}", flush=True)
    s = Stream(name, slots, live_threshold=live, reconv=reconv)
    got = s.run_all(ins)
    want, err = O.run(name, ins)
    assert err == 0
    for g, w in zip(got, want):
        assert np.array_equal(bits(g), bits(w)), name
    c = s.counters()
    n = len(ins[0])
    in_row = sum(np.dtype(d).itemsize for d in s.kernel.inputs)
    out_row = sum(np.dtype(d).itemsize for d in s.kernel.outputs) + 8  # + sequence number
    assert c["rows_in"] == n and c["rows_out"] == n
    assert c["bytes_h2d"] == n * in_row
    assert c["bytes_d2h"] == n * out_row
    s.close()


def test_exchange_moves_only_halted_rows(gpu):
    """An exchange takes min(halted, offered) rows: the rest of the offered
    input stays with the caller, live threads keep running."""
    x = workloads.collatz_inputs(4096, seed=23)
    s = Stream("collatz_u64", 1024, live_threshold=56)
    assert s.set_inputs([x]) == 1024
    s.step()
    live, halted = s.state()
    assert live + halted == 1024 and halted > 0
    before = s.counters()
    outs, ids = s.exchange([x[1024:]])
    assert len(ids) == min(halted, 4096 - 1024)
    after = s.counters()
    assert after["bytes_h2d"] - before["bytes_h2d"] == 8 * len(ids)
    assert after["bytes_d2h"] - before["bytes_d2h"] == 12 * len(ids)
    want = O.collatz_direct(x[ids.astype(np.int64)])
    assert np.array_equal(outs[0], want)
    # the new threads are numbered after the first 1024
    assert set(ids.tolist()) <= set(range(1024))
    live2, halted2 = s.state()
    assert live2 == live + len(ids) and halted2 == halted - len(ids)
    s.close()


def test_sort_reports_exchange_availability(gpu):
    x = workloads.fib_inputs(100, seed=24)
    s = Stream("fib_u64", 128, live_threshold=0)
    s.set_inputs([x])
    assert not s.sort()           # nothing halted yet
    s.step()                      # live threshold 0: every thread runs to HALT
    assert s.done
    assert s.sort()               # all halted, none live
    outs, ids = s.drain()
    assert sorted(ids.tolist()) == list(range(100))
    want, _ = O.run("fib_u64", [x])
    assert np.array_equal(outs[0], want[0][ids.astype(np.int64)])
    assert not s.sort() and s.state() == (0, 0)
    s.close()


def test_stream_reused_after_set_inputs(gpu):
    """set_inputs starts the stream over (the reference stream is reused
    across runs by its tests)."""
    s = Stream("collatz_u64", 256, live_threshold=32)
    for seed in (25, 26):
        x = workloads.collatz_inputs(3000, seed=seed)
        got = s.run_all([x])
        assert np.array_equal(got[0], O.collatz_direct(x))
    s.close()
