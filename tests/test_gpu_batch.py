"""Batched persistent launches (one launch, several kernels, uneven sizes)
and their HIP-graph replays, checked job by job against single-kernel runs
and the oracle."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def dev(a):
    return loom_amd.DeviceArray.from_numpy(a)


def host(cols):
    return [c.numpy() for c in cols]


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def build_mixed(b):
    jobs = []
    specs = [("smallpt9_f32", workloads.pt_inputs(40, 24, 4)),
             ("fib_u64", [workloads.fib_inputs(3000, seed=5)]),
             ("collatz_u64", [workloads.collatz_inputs(5000, seed=6)]),
             ("smallpt9_f32", workloads.pt_inputs(16, 20, 7)),
             ("fib_u64", [workloads.fib_inputs(1, seed=7)]),
             ("collatz_u64", [workloads.collatz_inputs(129, seed=8)])]
    for name, ins in specs:
        k = loom_amd.kernel(name)
        din = [dev(a) for a in ins]
        dout = [loom_amd.DeviceArray(len(ins[0]), dt) for dt in k.outputs]
        b.add(k, din, dout)
        jobs.append((name, ins, dout))
    return jobs


def check_jobs(jobs):
    for name, ins, dout in jobs:
        want, err = O.run(name, ins)
        assert err == 0
        for g, w in zip(host(dout), want):
            assert np.array_equal(bits(g), bits(w)), name


@pytest.mark.parametrize("reconv", [abi.RECONV_SUCCESSOR_MAJOR, abi.RECONV_TOPO,
                                    abi.RECONV_SWEEP])
@pytest.mark.parametrize("live", [0, 32, 48, 60])
def test_mixed_batch_matches_oracle(gpu, live, reconv):
    b = loom_amd.Batch("pt_int_mix")
    jobs = build_mixed(b)
    print(f"[gpu] batch pt_int_mix live={live} reconv={reconv} jobs={len(jobs)}", flush=True)
    r = b.run(loom_amd.make_params(mode=abi.MODE_HYBRID, live_threshold=live, reconv=reconv,
                                   max_steps=1 << 26))
    assert r.errors == 0
    check_jobs(jobs)


def test_subrange_jobs(gpu):
    """Jobs over [begin, end) of shared columns leave the rest untouched."""
    x = workloads.collatz_inputs(10000, seed=9)
    k = loom_amd.kernel("collatz_u64")
    din = [dev(x)]
    dout = loom_amd.DeviceArray(len(x), np.uint32)
    dout.zero()
    b = loom_amd.Batch("int_mix")
    b.add(k, din, [dout], 100, 4000)
    b.add(k, din, [dout], 6000, 6001)
    b.add(k, din, [dout], 7000, 7000)  # empty
    print("[gpu] batch int_mix subranges", flush=True)
    assert b.run().errors == 0
    (w,), _ = O.run("collatz_u64", [x])
    got = dout.numpy()
    sel = np.zeros(len(x), bool)
    sel[100:4000] = True
    sel[6000] = True
    assert np.array_equal(got[sel], w[sel])
    assert not got[~sel].any()


def test_graph_replays_match_eager(gpu):
    """A captured batch replays several times, with eager work in between,
    and every replay reproduces the eager result."""
    b = loom_amd.Batch("pt_int_mix")
    jobs = build_mixed(b)
    print("[gpu] batch eager", flush=True)
    b.run()
    eager = [host(d) for _, _, d in jobs]
    g = b.graph()
    for rep in range(3):
        for _, _, dout in jobs:
            for d in dout:
                d.zero()
        print(f"[gpu] graph replay {rep}", flush=True)
        r = g.launch()
        assert r.errors == 0
        for (name, _, dout), e in zip(jobs, eager):
            for x, y in zip(host(dout), e):
                assert np.array_equal(bits(x), bits(y)), (rep, name)
        # eager work between replays
        loom_amd.run("fib_u64", abi.MODE_COMPACT, [workloads.fib_inputs(1000)])
    g.destroy()
    check_jobs(jobs)
