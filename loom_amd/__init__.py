"""loom_amd — MI355X (gfx950) execution backend for Loom kernel_blank IR.

Python mirror of the reference backend interface (the C++ facade for C++
callers is include/loom_vm/vm/hip.hpp):

    f = loom_amd.compact_function("fib_u64")    # simd::compact_function<K>
    (c,) = f(n)                                  # one host column per input
    h = loom_amd.hybrid_function("fib_u64")     # simd::hybrid_function<K>
    (c,) = h(n)

Inputs are numpy arrays (host) or DeviceArray (HBM-resident columns);
outputs come back in input order.  Everything goes through the C ABI of
include/loom_amd.h; nothing here computes a kernel on the CPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi

TYPE_DTYPES = {
    0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64,
    4: np.int8, 5: np.int16, 6: np.int32, 7: np.int64,
    8: np.float32, 9: np.float64,
}

_initialised = False


def init(device: int = 0) -> None:
    """Select the GPU (reference: implicit in the CUDA backend)."""
    global _initialised
    if not _initialised:
        abi.check(abi.load().loom_init(device), "loom_init")
        _initialised = True


def device_info() -> dict:
    init()
    cus, ws, mem = C.c_int(), C.c_int(), C.c_size_t()
    name = C.create_string_buffer(64)
    abi.check(abi.load().loom_device_info(C.byref(cus), C.byref(ws), C.byref(mem), name, 64),
              "loom_device_info")
    return {"compute_units": cus.value, "wave_size": ws.value,
            "global_mem": mem.value, "arch": name.value.decode()}


@dataclass
class Kernel:
    """Summary of a lowered kernel (loom_kernel_info)."""
    index: int
    name: str
    hash: int
    inputs: list
    outputs: list
    n_blocks: int
    n_registers: int
    index_bytes: int
    state_size: int
    queue_row_bytes: int
    block_ops: list = field(default_factory=list)
    block_live_bytes: list = field(default_factory=list)
    block_flops: list = field(default_factory=list)


def kernel(name_or_index) -> Kernel:
    lib = abi.load()
    if isinstance(name_or_index, Kernel):
        return name_or_index
    if isinstance(name_or_index, str):
        idx = lib.loom_kernel_find(name_or_index.encode())
        if idx < 0:
            raise KeyError(f"no kernel named {name_or_index!r}")
    else:
        idx = int(name_or_index)
    info = abi.KernelInfo()
    abi.check(lib.loom_kernel_info_get(idx, C.byref(info)), "loom_kernel_info_get")
    nb = info.n_blocks
    return Kernel(
        index=idx, name=info.name.decode(), hash=info.hash,
        inputs=[TYPE_DTYPES[info.input_type[k]] for k in range(info.n_inputs)],
        outputs=[TYPE_DTYPES[info.output_type[k]] for k in range(info.n_outputs)],
        n_blocks=nb, n_registers=info.n_registers, index_bytes=info.index_bytes,
        state_size=info.state_size, queue_row_bytes=info.queue_row_bytes,
        block_ops=list(info.block_ops[:nb]), block_live_bytes=list(info.block_live_bytes[:nb]),
        block_flops=list(info.block_flops[:nb]))


def kernels() -> list[Kernel]:
    lib = abi.load()
    return [kernel(i) for i in range(lib.loom_kernel_count())]


class DeviceArray:
    """A 1-D column in device memory (owned)."""

    def __init__(self, n: int, dtype):
        init()
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        abi.check(abi.load().loom_malloc(C.byref(p), max(1, self.nbytes)), "loom_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            abi.check(abi.load().loom_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes), "h2d")
        return d

    def numpy(self, begin: int = 0, end: int | None = None) -> np.ndarray:
        """Elements [begin, end) copied to the host (the whole column by
        default)."""
        end = self.n if end is None else min(int(end), self.n)
        begin = min(max(0, int(begin)), end)
        out = np.empty(end - begin, self.dtype)
        if out.size:
            abi.check(abi.load().loom_memcpy_d2h(
                out.ctypes.data, self.ptr + begin * self.dtype.itemsize, out.nbytes), "d2h")
        return out

    def zero(self) -> None:
        abi.check(abi.load().loom_memset(self.ptr, 0, max(1, self.nbytes)), "memset")

    @classmethod
    def full(cls, n: int, value, dtype) -> "DeviceArray":
        """A column of n copies of `value`, filled on the device."""
        d = cls(n, dtype)
        v = np.array([value], dtype=d.dtype)
        abi.check(abi.load().loom_fill(d.ptr, v.ctypes.data, d.dtype.itemsize, d.n), "fill")
        return d

    def free(self) -> None:
        if self.ptr:
            abi.load().loom_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class RunResult:
    outputs: list
    ms_total: float
    ms_main: float
    main_launches: int
    passes: int
    errors: int
    grid_waves: int
    counters: list

    @property
    def active_lane_fraction(self) -> float:
        slot = self.counters[abi.C_SLOT_OPS]
        return self.counters[abi.C_ACTIVE_OPS] / slot if slot else float("nan")


def make_params(threshold=0, live_threshold=None, refill=True, regroup=False, waves_per_cu=0,
                queue_capacity=0, max_steps=0, stats=True, reconv=None,
                mode=abi.MODE_COMPACT, register_file=abi.REGFILE_VGPR,
                regroup_phase=0, regroup_waves=0) -> abi.Params:
    """Run parameters; `reconv` defaults to successor-major for the compact
    form (the reference policy) and to the sweep policy for hybrid,
    `live_threshold` to 48 for hybrid (0 for compact, which does not
    refill); `register_file` = REGFILE_MEMORY keeps a compact run's
    registers in device memory (sweep policy); `regroup` = REGROUP_WORKGROUP
    sorts a workgroup's threads by block between run phases of
    `regroup_phase` dispatches, in workgroups of `regroup_waves` waves (16,
    one per CU, or 8, two per CU; 0 = 16)."""
    p = abi.Params()
    if reconv is None:
        reconv = abi.RECONV_SUCCESSOR_MAJOR if mode == abi.MODE_COMPACT else abi.RECONV_SWEEP
    if live_threshold is None:
        live_threshold = 48 if mode == abi.MODE_HYBRID else 0
    p.reconv = int(reconv)
    p.threshold = int(threshold)
    p.live_threshold = int(live_threshold)
    p.refill = 1 if refill else 0
    p.regroup = int(regroup)  # False/True, or abi.REGROUP_*
    p.waves_per_cu = int(waves_per_cu)
    p.queue_capacity = int(queue_capacity)
    p.max_steps = int(max_steps)
    p.collect_stats = 1 if stats else 0
    p.register_file = int(register_file)
    p.regroup_phase = int(regroup_phase)
    p.regroup_waves = int(regroup_waves)
    return p


def run(k, mode: int, inputs, outputs=None, n=None, params: abi.Params | None = None,
        context: "Context | None" = None) -> RunResult:
    """Run every thread of `inputs` to HALT on the GPU.

    `inputs` are numpy arrays (copied in/out through the C ABI) or
    DeviceArray columns (then `outputs` must be DeviceArray columns too and
    the run reads and writes HBM only).  `context` is an execution context
    of the calling thread (default: the library's default context)."""
    init()
    k = kernel(k)
    lib = abi.load()
    if len(inputs) != len(k.inputs):
        raise ValueError(f"{k.name} takes {len(k.inputs)} inputs, got {len(inputs)}")
    params = params or make_params(mode=mode)
    st = abi.RunStats()
    ctx = context.handle if context is not None else None
    device = isinstance(inputs[0], DeviceArray)
    if n is None:
        n = inputs[0].n if device else len(inputs[0])
    if device:
        if outputs is None:
            outputs = [DeviceArray(n, dt) for dt in k.outputs]
        for a, dt in zip(inputs, k.inputs):
            if a.dtype != np.dtype(dt) or a.n < n:
                raise ValueError("device input column of wrong type or size")
        ins = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in inputs])
        outs = (C.c_void_p * abi.MAX_IO)(*[a.ptr for a in outputs])
        abi.check(lib.loom_ctx_run(ctx, k.index, mode, ins, outs, n, C.byref(params),
                                   C.byref(st)), "loom_run")
    else:
        host_in = [np.ascontiguousarray(a, dtype=dt) for a, dt in zip(inputs, k.inputs)]
        for a in host_in:
            if len(a) != n:
                raise ValueError("input columns differ in length")
        if outputs is None:
            outputs = [np.zeros(n, dt) for dt in k.outputs]
        ins = (C.c_void_p * abi.MAX_IO)(*[a.ctypes.data for a in host_in])
        outs = (C.c_void_p * abi.MAX_IO)(*[a.ctypes.data for a in outputs])
        abi.check(lib.loom_ctx_run_host(ctx, k.index, mode, ins, outs, n, C.byref(params),
                                        C.byref(st)), "loom_run_host")
    return RunResult(outputs=outputs, ms_total=st.ms_total, ms_main=st.ms_main,
                     main_launches=st.main_launches, passes=st.passes, errors=st.errors,
                     grid_waves=st.grid_waves, counters=list(st.counters))


class Context:
    """An execution context for one host thread (loom_context: a kernel
    stream and a copy stream of its own, its counters and workspaces).
    Threads with contexts of their own run concurrently and their kernels
    overlap on the GPU (the reference's Shop workers, src/vm/mp.hpp:158-190)."""

    def __init__(self):
        init()
        h = C.c_void_p()
        abi.check(abi.load().loom_context_create(C.byref(h)), "loom_context_create")
        self.handle = h

    def run(self, k, mode: int, inputs, outputs=None, n=None, params=None) -> RunResult:
        return run(k, mode, inputs, outputs, n, params, context=self)

    def close(self):
        if self.handle:
            abi.load().loom_context_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class compact_function:
    """Predicated wave64 form (reference simd::compact_function<K>,
    src/vm/simd.hpp:4329-4335): `f(*inputs)` returns the output columns."""

    mode = abi.MODE_COMPACT

    def __init__(self, k, **params):
        self.kernel = kernel(k)
        self.params = make_params(mode=self.mode, **params)
        self.last: RunResult | None = None

    def __call__(self, *inputs):
        self.last = run(self.kernel, self.mode, list(inputs), params=self.params)
        return tuple(self.last.outputs)


class hybrid_function(compact_function):
    """Regrouping persistent form (reference simd::hybrid_function<K>,
    src/vm/simd.hpp:4310-4317).  `threshold` is the regroup threshold in lanes
    (the reference UTRt·N, simd.hpp:4094-4098), `refill` admits new threads
    into free lanes when at most `live_threshold` lanes are live, `regroup`
    spills survivors to per-label queues once no input is left,
    `waves_per_cu` sizes the persistent grid."""

    mode = abi.MODE_HYBRID

    def __init__(self, k, threshold=0, live_threshold=48, refill=True, regroup=False,
                 waves_per_cu=0, queue_capacity=0, **params):
        super().__init__(k, threshold=threshold, live_threshold=live_threshold, refill=refill,
                         regroup=regroup, waves_per_cu=waves_per_cu,
                         queue_capacity=queue_capacity, **params)


# ---- batched launches (reference mp::Shop, src/vm/mp.hpp:316-385) ------------------
def _result(st, outputs=None):
    return RunResult(outputs=outputs, ms_total=st.ms_total, ms_main=st.ms_main,
                     main_launches=st.main_launches, passes=st.passes, errors=st.errors,
                     grid_waves=st.grid_waves, counters=list(st.counters))


class Batch:
    """Jobs of several kernels run by ONE persistent launch of a compiled
    batch set.  `add(kernel, inputs, outputs, begin, end)` takes DeviceArray
    columns; `run()` executes every job to completion; `graph()` captures the
    batch into a HIP graph for replay."""

    def __init__(self, set_name: str):
        init()
        lib = abi.load()
        self.set = lib.loom_batch_set_find(set_name.encode())
        if self.set < 0:
            raise KeyError(f"no batch set {set_name!r}")
        buf = (C.c_int * 64)()
        n = lib.loom_batch_set_members(self.set, buf, 64)
        self.members = [kernel(buf[i]) for i in range(n)]
        self.jobs = []
        self._keep = []

    def add(self, k, inputs, outputs, begin=0, end=None):
        k = kernel(k)
        if k.hash not in {m.hash for m in self.members}:
            raise ValueError(f"{k.name} is not a member of this batch set")
        if len(self.jobs) >= abi.MAX_JOBS:
            raise ValueError("a batch holds at most 64 jobs")
        j = abi.Job()
        j.kernel = k.index
        j.begin = int(begin)
        j.end = int(inputs[0].n if end is None else end)
        for i, a in enumerate(inputs):
            j.inputs[i] = a.ptr
        for i, a in enumerate(outputs):
            j.outputs[i] = a.ptr
        self.jobs.append(j)
        self._keep.append((inputs, outputs))
        return len(self.jobs) - 1

    def _table(self):
        return (abi.Job * len(self.jobs))(*self.jobs)

    def run(self, params=None) -> RunResult:
        params = params or make_params(mode=abi.MODE_HYBRID, live_threshold=48)
        st = abi.RunStats()
        abi.check(abi.load().loom_batch_run(self.set, self._table(), len(self.jobs),
                                            C.byref(params), C.byref(st)), "loom_batch_run")
        return _result(st)

    def graph(self, params=None) -> "Graph":
        params = params or make_params(mode=abi.MODE_HYBRID, live_threshold=48)
        g = C.c_void_p()
        abi.check(abi.load().loom_batch_graph_create(self.set, self._table(), len(self.jobs),
                                                     C.byref(params), C.byref(g)),
                  "loom_batch_graph_create")
        return Graph(g, self._keep)


class Graph:
    def __init__(self, handle, keep):
        self.handle = handle
        self._keep = keep

    def launch(self) -> RunResult:
        st = abi.RunStats()
        abi.check(abi.load().loom_graph_launch(self.handle, C.byref(st)), "loom_graph_launch")
        return _result(st)

    def destroy(self):
        if self.handle:
            abi.load().loom_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
