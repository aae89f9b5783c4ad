"""Serving: continuous batching of kernel launches (requests) on one GPU.

`Server` wraps loom_server_* (loom_amd/csrc/serve.cpp): requests are
admitted FIFO by free job slots and HBM budget, and every step is one batch
launch that advances all resident requests together.  `poisson_trace` draws
the mixed request stream (path-trace frames 128^2-512^2 at 4-16 spp, fib and
Collatz launches of 2^14-2^20 threads) with Poisson arrivals at a fixed rate,
all from splitmix64 so the trace is identical everywhere.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import abi, init, kernel, make_params, workloads

PT_MAX = 512 * 512 * 16
INT_MAX_LOG2 = 20


class Pinned:
    """A numpy array in pinned host memory (hipHostMalloc)."""

    def __init__(self, n: int, dtype):
        init()
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = C.c_void_p()
        abi.check(abi.load().loom_host_alloc(C.byref(p), max(1, self.n * self.dtype.itemsize)),
                  "loom_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(1, self.n * self.dtype.itemsize)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.n)

    def free(self):
        if self.ptr:
            self.array = None
            abi.load().loom_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class RequestSpec:
    kernel: str
    n: int
    arrival_ms: float
    w: int = 0
    h: int = 0
    spp: int = 0
    offset: int = 0  # fib/Collatz: first element of the shared input pool
    inputs: list | None = None  # host columns of any other kernel's request


def poisson_trace(qps: float, count: int, seed: int = workloads.SEED + 3,
                  pt_fraction: float = 0.5, scale: int = 1) -> list[RequestSpec]:
    """`count` requests with exponential inter-arrival gaps of mean 1/qps.
    Half path-trace frames (w, h uniform in [128, 512], spp in [4, 16]), the
    rest alternately fib and Collatz with 2^k threads, k uniform in
    [14, 20], their inputs a slice of a shared seeded pool.  `scale` > 1
    divides every size (for tests)."""
    r = workloads.splitmix64(seed, 6 * count)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    t = 0.0
    out = []
    n_int = 0
    for i in range(count):
        ua, ub, uc, ud, ue, uf = (float(x) for x in u[6 * i:6 * i + 6])
        t += -math.log(1.0 - ua) / qps * 1e3
        if ub < pt_fraction:
            w = max(1, (128 + int(uc * 385)) // scale)
            h = max(1, (128 + int(ud * 385)) // scale)
            spp = max(1, (4 + int(ue * 13)) // scale)
            out.append(RequestSpec("smallpt9_f32", w * h * spp, t, w, h, spp))
        else:
            n = max(1, (1 << (14 + int(uc * (INT_MAX_LOG2 - 13)))) // scale)
            name = "fib_u64" if n_int % 2 == 0 else "collatz_u64"
            n_int += 1
            off = int(uf * ((1 << (INT_MAX_LOG2 + 1)) - n))
            out.append(RequestSpec(name, n, t, offset=off))
    return out


@dataclass
class Pools:
    """Shared pinned input pools of a trace."""
    sid: Pinned = None
    fib: Pinned = None
    collatz: Pinned = None

    @classmethod
    def make(cls):
        p = cls()
        p.sid = Pinned(PT_MAX, np.uint32)
        p.sid.array[:] = np.arange(PT_MAX, dtype=np.uint32)
        n = 1 << (INT_MAX_LOG2 + 1)
        p.fib = Pinned(n, np.uint64)
        p.fib.array[:] = workloads.fib_inputs(n, seed=workloads.SEED + 4)
        p.collatz = Pinned(n, np.uint64)
        p.collatz.array[:] = workloads.collatz_inputs(n, seed=workloads.SEED + 5)
        return p

    def inputs_of(self, spec: RequestSpec, full: bool = False):
        """Host input columns of a request (views of the pools, or the
        request's own columns).  A frame's second input is one word
        broadcast to every thread; `full` spells it out as a column."""
        if spec.inputs is not None:
            return spec.inputs
        if spec.kernel.startswith("smallpt9"):
            job = workloads.pt_job(spec.w, spec.h, spec.spp)
            return [self.sid.array[:spec.n],
                    np.full(spec.n, job, np.uint32) if full else None]
        pool = self.fib if spec.kernel.startswith("fib") else self.collatz
        return [pool.array[spec.offset:spec.offset + spec.n]]


class Server:
    """Continuous-batching server over a compiled batch set."""

    def __init__(self, set_name: str = "pt_int_mix", hbm_budget: int = 8 << 30,
                 max_resident: int = 32, step_threads: int = 4 << 20, **params):
        init()
        lib = abi.load()
        cfg = abi.ServerConfig()
        cfg.batch_set = lib.loom_batch_set_find(set_name.encode())
        if cfg.batch_set < 0:
            raise KeyError(f"no batch set {set_name!r}")
        cfg.max_resident = int(max_resident)
        cfg.hbm_budget = int(hbm_budget)
        cfg.step_threads = int(step_threads)
        params.setdefault("live_threshold", 48)
        cfg.params = make_params(mode=abi.MODE_HYBRID, **params)
        h = C.c_void_p()
        abi.check(lib.loom_server_create(C.byref(cfg), C.byref(h)), "loom_server_create")
        self.handle = h
        self.config = cfg

    def run(self, specs: list[RequestSpec], pools: Pools, realtime: bool = True,
            outputs: list | None = None):
        """Serve `specs`; returns (requests array, stats, outputs).  Without
        `outputs`, every request gets its own output columns."""
        reqs = (abi.Request * len(specs))()
        keep = []
        outs = outputs if outputs is not None else []
        for i, s in enumerate(specs):
            k = kernel(s.kernel)
            q = reqs[i]
            q.kernel = k.index
            q.n = s.n
            q.arrival_ms = s.arrival_ms
            ins = pools.inputs_of(s)
            if s.inputs is None and s.kernel.startswith("smallpt9"):
                # the frame word is the same for every thread: broadcast it
                q.inputs[0] = ins[0].ctypes.data
                q.broadcast = 1 << 1
                q.scalars[1] = workloads.pt_job(s.w, s.h, s.spp)
            else:
                for j, col in enumerate(ins):
                    q.inputs[j] = col.ctypes.data
            keep.append(ins)
            if outputs is None:
                outs.append([np.zeros(s.n, dt) for dt in k.outputs])
            for j, o in enumerate(outs[i]):
                q.outputs[j] = o.ctypes.data
        st = abi.ServerStats()
        abi.check(abi.load().loom_server_run(self.handle, reqs, len(specs), 1 if realtime else 0,
                                             C.byref(st)), "loom_server_run")
        return reqs, st, outs

    # ---- online serving: the loop on a thread of its own, submissions from
    # any thread (loom_server_start / submit / wait / stop) ----
    def start(self):
        abi.check(abi.load().loom_server_start(self.handle), "loom_server_start")

    def submit(self, spec: RequestSpec, pools: Pools, outputs=None) -> "Ticket":
        """Submit one request while the loop runs (thread-safe); its arrival
        is now.  The returned ticket keeps the request and its columns alive
        until it is complete."""
        k = kernel(spec.kernel)
        q = abi.Request()
        q.kernel = k.index
        q.n = spec.n
        ins = pools.inputs_of(spec)
        q.inputs[0] = ins[0].ctypes.data
        if spec.kernel.startswith("smallpt9"):
            q.broadcast = 1 << 1
            q.scalars[1] = workloads.pt_job(spec.w, spec.h, spec.spp)
        outs = outputs if outputs is not None else [np.zeros(spec.n, dt) for dt in k.outputs]
        for j, o in enumerate(outs):
            q.outputs[j] = o.ctypes.data
        t = Ticket(q, ins, outs, self)
        abi.check(abi.load().loom_server_submit(self.handle, C.byref(q)), "loom_server_submit")
        return t

    def stop(self):
        st = abi.ServerStats()
        abi.check(abi.load().loom_server_stop(self.handle, C.byref(st)), "loom_server_stop")
        return st

    def close(self):
        if self.handle:
            abi.load().loom_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Ticket:
    """A submitted request: the C struct the server fills, and the host
    columns it reads and writes."""
    request: abi.Request
    inputs: list
    outputs: list
    server: Server

    def done(self) -> bool:
        d = C.c_int()
        abi.check(abi.load().loom_server_poll(self.server.handle, C.byref(self.request),
                                              C.byref(d)), "loom_server_poll")
        return bool(d.value)

    def wait(self) -> int:
        """Block until complete; returns the request's status."""
        return abi.load().loom_server_wait(self.server.handle, C.byref(self.request))


def latency_summary(reqs, st) -> dict:
    lat = np.array([q.done_ms - q.arrival_ms for q in reqs])
    wait = np.array([q.admit_ms - q.arrival_ms for q in reqs])
    return {
        "requests": len(reqs),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "mean_admit_wait_ms": float(wait.mean()),
        "makespan_ms": st.makespan_ms,
        "gpu_ms": st.gpu_ms,
        "gpu_busy": st.gpu_ms / st.makespan_ms if st.makespan_ms else 0.0,
        "steps": int(st.steps),
        "threads": int(st.threads),
        "threads_per_s": st.threads / (st.makespan_ms / 1e3) if st.makespan_ms else 0.0,
        "requests_per_s": len(reqs) / (st.makespan_ms / 1e3) if st.makespan_ms else 0.0,
        "peak_resident": int(st.peak_resident),
        "peak_bytes": int(st.peak_bytes),
    }
