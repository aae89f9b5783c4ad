"""The mixed batch: 64 seeded jobs of uneven sizes and different kernels
(path-trace frames, fib, Collatz) resident in HBM, run either as ONE
persistent batch launch (optionally replayed from a HIP graph) or one
launch per job, for timing and for job-by-job comparison."""
from __future__ import annotations

import time

import numpy as np

from . import Batch, DeviceArray, abi, autotune, init, kernel, make_params, run, workloads

SET = "pt_int_mix"


class MixedBatch:
    def __init__(self, jobs=None, scale: int = 1):
        init()
        self.specs = jobs if jobs is not None else workloads.batch_jobs(scale=scale)
        pt_max = max((s["n"] for name, s in self.specs if name.startswith("smallpt9")), default=0)
        # one shared sample-index column serves every frame (read-only input)
        self.sid = DeviceArray.from_numpy(np.arange(max(pt_max, 1), dtype=np.uint32))
        self.jobs = []
        for name, s in self.specs:
            k = kernel(name)
            n = s["n"]
            if name.startswith("smallpt9"):
                ins = [self.sid, DeviceArray.full(n, workloads.pt_job(s["w"], s["h"], s["spp"]),
                                                  np.uint32)]
            else:
                ins = [DeviceArray.from_numpy(workloads.job_inputs(name, s)[0])]
            outs = [DeviceArray(n, dt) for dt in k.outputs]
            self.jobs.append((k, n, ins, outs))
        self.threads = sum(n for _, n, _, _ in self.jobs)

    def batch(self) -> Batch:
        b = Batch(SET)
        for k, n, ins, outs in self.jobs:
            b.add(k, ins, outs, 0, n)
        return b

    def outputs(self):
        return [[o.numpy() for o in outs] for _, _, _, outs in self.jobs]

    def run_separately(self, params_of=None) -> float:
        """Each job as its own hybrid run (tuned per kernel); wall ms."""
        t = time.perf_counter()
        for k, n, ins, outs in self.jobs:
            p = params_of(k.name) if params_of else make_params(
                mode=abi.MODE_HYBRID, stats=False, **autotune.best_params(k.name))
            run(k, abi.MODE_HYBRID, ins, outs, n=n, params=p)
        abi.check(abi.load().loom_synchronize(), "sync")
        return (time.perf_counter() - t) * 1e3
