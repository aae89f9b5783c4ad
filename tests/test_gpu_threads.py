"""Several host threads feeding one GPU (the reference's Shop workers,
src/vm/mp.hpp:61-385): each thread runs kernels on an execution context of
its own, concurrently, and every result equals the same launch run alone;
threads sharing the default context serialise safely; the online server
takes submissions from several threads while its loop runs."""
import threading

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, serve, workloads
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def jobs_of(t):
    """A mixed list of launches for thread t."""
    return [("collatz_u64", abi.MODE_HYBRID, [workloads.collatz_inputs(30000 + 7 * t, seed=40 + t)]),
            ("smallpt9_f32", abi.MODE_HYBRID, workloads.pt_inputs(16 + t, 12, 4)),
            ("fib_u64", abi.MODE_COMPACT, [workloads.fib_inputs(5000 + t, seed=50 + t)]),
            ("cubic_bisect_f64", abi.MODE_HYBRID, workloads.cubic_inputs(2000 + t, np.float64))]


def run_threads(target, n):
    errs = []

    def wrap(t):
        try:
            target(t)
        except BaseException as e:  # reported by the main thread
            errs.append((t, e))

    ts = [threading.Thread(target=wrap, args=(t,)) for t in range(n)]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    assert not errs, errs


def test_threads_with_own_contexts_equal_alone(gpu):
    n = 4
    alone = {t: [loom_amd.run(k, m, ins).outputs for k, m, ins in jobs_of(t)] for t in range(n)}
    got = {}

    def work(t):
        with loom_amd.Context() as ctx:
            res = []
            for rep in range(3):
                res = [ctx.run(k, m, ins).outputs for k, m, ins in jobs_of(t)]
            got[t] = res

    print(f"[gpu] {n} threads, own contexts", flush=True)
    run_threads(work, n)
    for t in range(n):
        for (k, _, _), g, a in zip(jobs_of(t), got[t], alone[t]):
            for x, y in zip(g, a):
                assert np.array_equal(bits(x), bits(y)), (t, k)


def test_threads_sharing_the_default_context(gpu):
    """The entry points without a context serialise on the default
    context's lock: concurrent callers each get their own correct result."""
    n = 4
    got = {}

    def work(t):
        x = workloads.collatz_inputs(20000, seed=60 + t)
        for _ in range(3):
            got[t] = (x, loom_amd.run("collatz_u64", abi.MODE_HYBRID, [x]).outputs[0])

    print(f"[gpu] {n} threads, default context", flush=True)
    run_threads(work, n)
    for t in range(n):
        x, y = got[t]
        assert np.array_equal(y, O.collatz_direct(x)), t


def test_threads_submitting_to_online_server(gpu):
    """4 host threads submit mixed requests while the serving loop runs;
    every result equals the same launch alone."""
    pools = serve.Pools.make()
    specs = serve.poisson_trace(qps=1.0, count=40, scale=4, seed=13)
    srv = serve.Server(hbm_budget=64 << 20, max_resident=8, step_threads=1 << 16)
    srv.start()
    tickets = {}

    def work(t):
        mine = specs[t::4]
        tk = [srv.submit(s, pools) for s in mine]
        for k in tk:
            assert k.wait() == 0
        tickets[t] = list(zip(mine, tk))

    print("[gpu] online server, 4 submitting threads", flush=True)
    run_threads(work, 4)
    st = srv.stop()
    srv.close()
    assert st.threads == sum(s.n for s in specs)
    done = [p for t in range(4) for p in tickets[t]]
    assert len(done) == len(specs)
    for s, k in done:
        q = k.request
        assert q.complete == 1 and q.status == 0
        assert q.done_ms >= q.admit_ms >= q.arrival_ms - 1e-6
        alone = loom_amd.run(s.kernel, abi.MODE_HYBRID, pools.inputs_of(s, full=True)).outputs
        for g, a in zip(k.outputs, alone):
            assert np.array_equal(bits(g), bits(a)), (s.kernel, s.n)


def test_online_server_rejects_after_stop(gpu):
    pools = serve.Pools.make()
    srv = serve.Server(hbm_budget=16 << 20, max_resident=4, step_threads=1 << 14)
    srv.start()
    t = srv.submit(serve.RequestSpec("fib_u64", 1000, 0.0, offset=5), pools)
    assert t.wait() == 0
    srv.stop()
    with pytest.raises(abi.LoomError):
        srv.submit(serve.RequestSpec("fib_u64", 10, 0.0), pools)
    # and it can serve again after a restart
    srv.start()
    t = srv.submit(serve.RequestSpec("collatz_u64", 2000, 0.0, offset=9), pools)
    assert t.wait() == 0
    srv.stop()
    srv.close()
    want = O.collatz_direct(pools.collatz.array[9:2009])
    assert np.array_equal(t.outputs[0], want)
