"""Serving: a mixed Poisson request stream served with continuous batching
gives every request the same bits as the same launch run alone, while
requests share steps on the device; admission respects the HBM budget."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi, serve, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


@pytest.fixture(scope="module")
def pools(gpu):
    return serve.Pools.make()


def test_mixed_stream_equals_alone(pools):
    specs = serve.poisson_trace(qps=4000.0, count=48, scale=4, seed=11)
    srv = serve.Server(hbm_budget=48 << 20, max_resident=8, step_threads=1 << 16)
    print(f"[gpu] serve {len(specs)} requests realtime", flush=True)
    reqs, st, outs = srv.run(specs, pools, realtime=True)
    srv.close()
    assert all(q.status == 0 for q in reqs)
    # the mix really overlapped: several requests resident in one step,
    # and requests spread over several steps
    assert st.peak_resident > 1
    assert max(q.steps for q in reqs) > 1
    assert st.peak_bytes <= 48 << 20
    for s, q, o in zip(specs, reqs, outs):
        assert q.done_ms >= q.admit_ms >= q.arrival_ms - 1e-6
        alone = loom_amd.run(s.kernel, abi.MODE_HYBRID, pools.inputs_of(s, full=True)).outputs
        for g, a in zip(o, alone):
            assert np.array_equal(bits(g), bits(a)), (s.kernel, s.n)
    # and the oracle agrees on a few
    for s, o in list(zip(specs, outs))[:6]:
        want, err = O.run(s.kernel, pools.inputs_of(s, full=True))
        assert err == 0
        for g, w in zip(o, want):
            assert np.array_equal(bits(g), bits(w)), (s.kernel, s.n)
    summ = serve.latency_summary(reqs, st)
    print(summ)
    assert summ["p99_ms"] >= summ["p50_ms"] > 0


def test_budget_rejects_oversized_request(pools):
    specs = [serve.RequestSpec("collatz_u64", 5000, 0.0, offset=7),
             serve.RequestSpec("collatz_u64", 1 << 20, 0.0, offset=0),  # 12 MiB > budget
             serve.RequestSpec("fib_u64", 3000, 0.0, offset=100)]
    srv = serve.Server(hbm_budget=1 << 20, max_resident=4, step_threads=1 << 12)
    print("[gpu] serve budget rejection", flush=True)
    reqs, st, outs = srv.run(specs, pools, realtime=False)
    srv.close()
    assert [q.status for q in reqs] == [0, abi.E_CAPACITY, 0]
    for i in (0, 2):
        want, _ = O.run(specs[i].kernel, pools.inputs_of(specs[i]))
        assert np.array_equal(outs[i][0], want[0])


def test_kernels_outside_the_batch_set(pools):
    """Requests of kernels outside the server's batch set run their slice of
    each step as launches of their own, interleaved with the batch launches
    of the others; every result equals its run alone."""
    base = serve.poisson_trace(qps=4000.0, count=24, scale=4, seed=12)
    extra = []
    for i, (name, n) in enumerate((("cubic_bisect_f32", 70000), ("threefry2x32_u32", 50000),
                                   ("cubic_bisect_f64", 30000), ("collatz_u64", 9000))):
        if name.startswith("cubic"):
            ins = workloads.cubic_inputs(n, np.float32 if name.endswith("f32") else np.float64)
        elif name.startswith("threefry"):
            k = loom_amd.kernel(name)
            rng = np.random.default_rng(70 + i)
            ins = [rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(dt) for dt in k.inputs]
        else:
            ins = [workloads.collatz_inputs(n, seed=71)]
        extra.append(serve.RequestSpec(name, n, base[2 * i].arrival_ms + 0.01, inputs=ins))
    specs = sorted(base + extra, key=lambda s: s.arrival_ms)
    srv = serve.Server(hbm_budget=96 << 20, max_resident=8, step_threads=1 << 16)
    print(f"[gpu] serve {len(specs)} requests, 3 outside the batch set", flush=True)
    reqs, st, outs = srv.run(specs, pools, realtime=True)
    srv.close()
    assert all(q.status == 0 for q in reqs)
    assert max(q.steps for q in reqs) > 1
    for s, q, o in zip(specs, reqs, outs):
        alone = loom_amd.run(s.kernel, abi.MODE_HYBRID, pools.inputs_of(s, full=True)).outputs
        for g, a in zip(o, alone):
            assert np.array_equal(bits(g), bits(a)), (s.kernel, s.n)
