"""Kernel inspector: the lowered listing of a kernel (blocks, cost, live-in
registers, successors, instructions), its compiled resources, the device
counters of one run in each form (--run), and a step trace of one thread
recorded on the GPU (--trace), printed as the reference's step-by-step
inspector prints it: one row per block, the block label and the tracked
registers (src/vm/dbg.hpp:1099-1140, rows as print_state 920-935).

    python -m loom_amd.inspector smallpt9_f32            # listing (no GPU needed)
    python -m loom_amd.inspector collatz_u64 --run 65536 # + resources and counters
    python -m loom_amd.inspector fib_u64 --trace 10 --steps 40 --track n,c
"""
from __future__ import annotations

import argparse
import ctypes as C
import re
import sys
from dataclasses import dataclass

import numpy as np

# loom_type codes -> numpy types (include/loom_amd.h)
_DTYPES = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64, 4: np.int8, 5: np.int16,
           6: np.int32, 7: np.int64, 8: np.float32, 9: np.float64}


def listing(name) -> str:
    """The lowered listing of kernel `name` (or index)."""
    from loom_amd import abi, kernel
    k = kernel(name)
    lib = abi.load()
    need = C.c_size_t()
    abi.check(lib.loom_kernel_dump(k.index, None, 0, C.byref(need)), "loom_kernel_dump")
    buf = C.create_string_buffer(need.value)
    abi.check(lib.loom_kernel_dump(k.index, buf, need.value, None), "loom_kernel_dump")
    return buf.value.decode()


def registers(name) -> list[dict]:
    """Name, first type, width, type count and input/output/written flags of
    every register (no GPU)."""
    from loom_amd import abi, kernel
    k = kernel(name)
    lib = abi.load()
    out = []
    buf = C.create_string_buffer(256)
    for r in range(k.n_registers):
        ty, w, nt, fl = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        abi.check(lib.loom_kernel_register(k.index, r, buf, 256, C.byref(ty), C.byref(w),
                                           C.byref(nt), C.byref(fl)), "loom_kernel_register")
        out.append({"name": buf.value.decode(), "dtype": np.dtype(_DTYPES[ty.value]),
                    "width": w.value, "ntypes": nt.value,
                    "input": bool(fl.value & abi.REG_INPUT),
                    "output": bool(fl.value & abi.REG_OUTPUT),
                    "written": bool(fl.value & abi.REG_WRITTEN)})
    return out


def block_labels(name) -> list[str]:
    """Label of every block, in block order, as the listing names them."""
    return [m.group(2) for m in re.finditer(r"^block (\d+) (\S+)", listing(name), re.M)]


@dataclass
class Trace:
    """Step trace of threads [0, n): per thread, the block run at each step,
    the next block, the error bits and the raw bits of every register after
    the block (uint64, zero-extended)."""
    kernel: str
    registers: list
    labels: list
    steps: np.ndarray        # [n] rows per thread
    block: np.ndarray        # [n, max_steps] uint32
    next_block: np.ndarray   # [n, max_steps] uint32 (abi.TRACE_HALT after the last block)
    errors: np.ndarray       # [n, max_steps] uint64
    bits: np.ndarray         # [n, max_steps, n_registers] uint64

    def value(self, thread: int, step: int, reg):
        """Register `reg` (name or index) of `thread` after `step`, as its type."""
        r = reg if isinstance(reg, int) else [g["name"] for g in self.registers].index(reg)
        dt = self.registers[r]["dtype"]
        return np.array([self.bits[thread, step, r]], np.uint64).view(np.uint8)[
            :dt.itemsize].view(dt)[0]


def trace(name, inputs, max_steps: int = 100) -> Trace:
    """Run threads [0, len(inputs[0])) one block at a time on the GPU
    (loom_kernel_trace) and return their rows."""
    import loom_amd
    from loom_amd import abi
    loom_amd.init()
    k = loom_amd.kernel(name)
    if len(inputs) != len(k.inputs):
        raise ValueError(f"{k.name} takes {len(k.inputs)} inputs, got {len(inputs)}")
    cols = [np.ascontiguousarray(a, dtype=dt) for a, dt in zip(inputs, k.inputs)]
    n = len(cols[0])
    if any(len(c) != n for c in cols):
        raise ValueError("input columns differ in length")
    words = abi.trace_words(k.n_registers)
    dev = [loom_amd.DeviceArray.from_numpy(c) for c in cols]
    rows = loom_amd.DeviceArray(max(1, n * max_steps * words), np.uint64)
    steps = loom_amd.DeviceArray(max(1, n), np.uint32)
    ins = (C.c_void_p * abi.MAX_IO)(*[d.ptr for d in dev])
    abi.check(abi.load().loom_kernel_trace(k.index, ins, n, max_steps, rows.ptr, rows.nbytes,
                                           steps.ptr), "loom_kernel_trace")
    R = rows.numpy()[:n * max_steps * words].reshape(n, max_steps, words)
    return Trace(kernel=k.name, registers=registers(k), labels=block_labels(k),
                 steps=steps.numpy()[:n].copy(), block=(R[:, :, 0] & 0xFFFFFFFF).astype(np.uint32),
                 next_block=(R[:, :, 0] >> np.uint64(32)).astype(np.uint32),
                 errors=R[:, :, 1].copy(), bits=R[:, :, 2:].copy())


def format_trace(tr: Trace, thread: int = 0, track=None, width: int = 20) -> str:
    """The rows of one thread as a table: BLOCK, then the tracked registers
    (all when `track` is empty), one row per block run."""
    names = [g["name"] for g in tr.registers]
    cols = [names.index(t) for t in track] if track else list(range(len(names)))

    def cell(text):
        text = str(text)
        return (text if len(text) < width else text[:width - 2] + "~").ljust(width - 1) + "|"

    lw = max([5] + [len(l) for l in tr.labels]) + 1
    out = ["BLOCK".ljust(lw) + "|" + "".join(cell(names[c]) for c in cols),
           "-" * lw + "|" + "".join("-" * (width - 1) + "|" for _ in cols)]
    for s in range(int(tr.steps[thread])):
        vals = []
        for c in cols:
            v = tr.value(thread, s, c)
            vals.append(f"0x{int(tr.bits[thread, s, c]):x}" if tr.registers[c]["ntypes"] > 1
                        else (repr(float(v)) if np.dtype(type(v)).kind == "f" else int(v)))
        out.append(tr.labels[int(tr.block[thread, s])].ljust(lw) + "|" +
                   "".join(cell(v) for v in vals))
    from loom_amd import abi
    n = int(tr.steps[thread])
    if n and int(tr.next_block[thread, n - 1]) == abi.TRACE_HALT:
        out.append("HALT")
    return "\n".join(out)


def resources(name) -> dict:
    """VGPRs and scratch bytes of every compiled form and policy."""
    import loom_amd
    from loom_amd import abi
    loom_amd.init()
    k = loom_amd.kernel(name)
    out = {}
    forms = ((abi.MODE_COMPACT, "compact", 0), (abi.MODE_HYBRID, "hybrid", 0),
             (abi.MODE_HYBRID, "hybrid-regroup", abi.RESOURCES_REGROUP))
    for mode, mname, flag in forms:
        for pol, pname in ((0, "successor-major"), (1, "topo"), (2, "sweep")):
            v, sc = C.c_int(), C.c_int()
            abi.check(abi.load().loom_kernel_resources(k.index, mode, pol | flag, C.byref(v),
                                                       C.byref(sc)), "loom_kernel_resources")
            out[f"{mname}/{pname}"] = {"vgprs": v.value, "scratch_bytes": sc.value}
    # the compact kernel with its register file in device memory (sweep)
    v, sc = C.c_int(), C.c_int()
    abi.check(abi.load().loom_kernel_resources(k.index, abi.MODE_COMPACT,
                                               abi.RECONV_SWEEP | abi.RESOURCES_MEMORY,
                                               C.byref(v), C.byref(sc)), "loom_kernel_resources")
    out["compact-memory/sweep"] = {"vgprs": v.value, "scratch_bytes": sc.value}
    # the workgroup-regrouping kernel (successor-major, sweep; 16- and
    # 8-wave workgroups)
    for flag, wname in ((abi.RESOURCES_WORKGROUP, "hybrid-workgroup"),
                        (abi.RESOURCES_WORKGROUP_8, "hybrid-workgroup8")):
        for pol, pname in ((0, "successor-major"), (2, "sweep")):
            v, sc = C.c_int(), C.c_int()
            abi.check(abi.load().loom_kernel_resources(k.index, abi.MODE_HYBRID, pol | flag,
                                                       C.byref(v), C.byref(sc)),
                      "loom_kernel_resources")
            out[f"{wname}/{pname}"] = {"vgprs": v.value, "scratch_bytes": sc.value}
    return out


def counters(name, n: int) -> dict:
    """Counters of one run per form on seeded inputs of n threads."""
    import numpy as np
    import loom_amd
    from loom_amd import abi, workloads
    loom_amd.init()
    k = loom_amd.kernel(name)
    ins = []
    for j, dt in enumerate(k.inputs):
        raw = workloads.splitmix64(workloads.SEED + 40 + j, n)
        dt = np.dtype(dt)
        ins.append(workloads.uniform_real(workloads.SEED + 40 + j, n, -8.0, 8.0, dt)
                   if dt.kind == "f" else (raw % np.uint64(64)).astype(dt))
    names = ["active_ops", "slot_ops", "dispatches", "thread_steps", "spill_rows", "spill_bytes",
             "fill_rows", "fill_bytes", "admitted", "refills", "batches", "spill_failed",
             "active_flops"]
    out = {}
    for mode, mname in ((abi.MODE_COMPACT, "compact"), (abi.MODE_HYBRID, "hybrid")):
        r = loom_amd.run(k, mode, ins, params=loom_amd.make_params(mode=mode))
        out[mname] = {"ms": r.ms_main, "active_lane_fraction": r.active_lane_fraction,
                      **{nm: int(v) for nm, v in zip(names, r.counters)}}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel")
    ap.add_argument("--run", type=int, default=0, help="threads of a counter run (needs a GPU)")
    ap.add_argument("--trace", default=None,
                    help="comma-separated input values of one thread to trace (needs a GPU)")
    ap.add_argument("--steps", type=int, default=100, help="maximum blocks traced (default 100)")
    ap.add_argument("--track", default="", help="comma-separated registers to show (default all)")
    ap.add_argument("--width", type=int, default=20, help="column width (default 20)")
    a = ap.parse_args(argv)
    if a.trace is not None:
        import loom_amd
        k = loom_amd.kernel(a.kernel)
        vals = a.trace.split(",")
        if len(vals) != len(k.inputs):
            ap.error(f"{k.name} takes {len(k.inputs)} inputs")
        cols = [np.array([float(v) if np.dtype(dt).kind == "f" else int(v, 0)], dtype=dt)
                for v, dt in zip(vals, k.inputs)]
        tr = trace(k, cols, a.steps)
        print(format_trace(tr, 0, [t for t in a.track.split(",") if t], a.width))
        return 0
    print(listing(a.kernel))
    if a.run:
        import json
        print(json.dumps({"resources": resources(a.kernel), "counters": counters(a.kernel, a.run)},
                         indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
