"""The 64-job mixed batch workload (path-trace frames with w, h uniform in
[64, 1024] and spp in [1, 64], fib and Collatz jobs of 2^12-2^22 threads,
drawn from a fixed seed) as ONE persistent launch replayed from a HIP graph
with the tuned parameters: every job checked against the oracle at reduced
scale, and at full scale fib against its closed form, Collatz against the
definition iterated directly (brute force) and a sampled prefix plus a
seeded sample of every frame against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, autotune, mixed, workloads
from tests import oracle_lib as O

CPU_THREADS = 16  # the host share of one GPU on the test box


def bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def fib_table():
    f = [0, 1]
    while len(f) < 91:
        f.append(f[-1] + f[-2])
    return np.array(f, dtype=np.uint64)  # F(90) < 2^64


def tuned_params():
    bp = autotune.best_params("batch:" + mixed.SET)
    return loom_amd.make_params(mode=abi.MODE_HYBRID, stats=True, **bp)


def replay(mb, params, times=2):
    """Replays the captured batch `times` times, clearing the outputs
    between replays; returns the last run's result."""
    g = mb.batch().graph(params)
    r = None
    for rep in range(times):
        for _, _, _, outs in mb.jobs:
            for o in outs:
                o.zero()
        print(f"[gpu] C4 graph replay {rep}: {len(mb.jobs)} jobs, {mb.threads} threads",
              flush=True)
        r = g.launch()
        assert r.errors == 0
    g.destroy()
    return r


def test_c4_table_shape():
    """The workload the batch runs: 64 jobs, half frames, a quarter each fib
    and Collatz, sizes in the stated ranges."""
    specs = workloads.batch_jobs()
    assert len(specs) == 64
    names = [n for n, _ in specs]
    assert names.count("smallpt9_f32") == 32
    assert names.count("fib_u64") == 16 and names.count("collatz_u64") == 16
    for name, s in specs:
        if name == "smallpt9_f32":
            assert 64 <= s["w"] <= 1024 and 64 <= s["h"] <= 1024 and 1 <= s["spp"] <= 64
        else:
            assert (1 << 12) <= s["n"] <= (1 << 22)


@pytest.mark.gpu
def test_c4_scaled_batch_matches_oracle(gpu):
    """Every job of the batch at 1/8 of the sizes, every thread against the
    oracle, after two graph replays."""
    mb = mixed.MixedBatch(scale=8)
    r = replay(mb, tuned_params())
    assert r.counters[abi.C_THREAD_STEPS] > 0
    for j, ((name, spec), (k, n, ins, outs)) in enumerate(zip(mb.specs, mb.jobs)):
        host_in = workloads.job_inputs(name, spec)
        want, err = O.run(name, host_in, threads=CPU_THREADS)
        assert err == 0
        for c, (o, w) in enumerate(zip(outs, want)):
            got = o.numpy()
            assert np.array_equal(bits(got), bits(w)), (j, name, spec, c)


@pytest.mark.gpu
def test_c4_full_batch(gpu):
    """The full-size 64-job batch (316,853,173 threads) in one graph-replayed
    launch: fib jobs against F(n), Collatz jobs against brute force, and for
    every frame its first 2048 samples and 2048 seeded samples against the
    oracle."""
    mb = mixed.MixedBatch()
    assert mb.threads == 316_853_173
    replay(mb, tuned_params())
    F = fib_table()
    rng = np.random.default_rng(0xC4)
    for j, ((name, spec), (k, n, ins, outs)) in enumerate(zip(mb.specs, mb.jobs)):
        if name == "fib_u64":
            x = workloads.job_inputs(name, spec)[0]
            assert np.array_equal(outs[0].numpy(), F[x]), (j, spec)
        elif name == "collatz_u64":
            x = workloads.job_inputs(name, spec)[0]
            want = O.collatz_direct(x, threads=CPU_THREADS)
            assert np.array_equal(outs[0].numpy(), want), (j, spec)
        else:
            head = np.arange(min(n, 2048), dtype=np.uint32)
            pick = np.unique(np.concatenate([head, rng.integers(0, n, 2048).astype(np.uint32)]))
            job = np.full(pick.size, workloads.pt_job(spec["w"], spec["h"], spec["spp"]),
                          np.uint32)
            want, err = O.run(name, [pick, job], threads=CPU_THREADS)
            assert err == 0
            for c, (o, w) in enumerate(zip(outs, want)):
                got = o.numpy()[pick]
                assert np.array_equal(bits(got), bits(w)), (j, spec, c)
