"""Kernel inspector: the lowered listing of a kernel (blocks, cost, live-in
registers, successors, instructions), its compiled resources and, with
--run, the device counters of one run in each form.

    python -m loom_amd.inspect smallpt9_f32            # listing (no GPU needed)
    python -m loom_amd.inspect collatz_u64 --run 65536 # + resources and counters
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys


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


def resources(name) -> dict:
    """VGPRs and scratch bytes of every compiled form and policy."""
    import loom_amd
    from loom_amd import abi
    loom_amd.init()
    k = loom_amd.kernel(name)
    out = {}
    for mode, mname in ((abi.MODE_COMPACT, "compact"), (abi.MODE_HYBRID, "hybrid")):
        for pol, pname in ((0, "successor-major"), (1, "topo"), (2, "sweep")):
            v, sc = C.c_int(), C.c_int()
            abi.check(abi.load().loom_kernel_resources(k.index, mode, pol, C.byref(v),
                                                       C.byref(sc)), "loom_kernel_resources")
            out[f"{mname}/{pname}"] = {"vgprs": v.value, "scratch_bytes": sc.value}
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
    a = ap.parse_args(argv)
    print(listing(a.kernel))
    if a.run:
        import json
        print(json.dumps({"resources": resources(a.kernel), "counters": counters(a.kernel, a.run)},
                         indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
