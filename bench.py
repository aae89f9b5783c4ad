#!/usr/bin/env python3
"""Benchmark: path-traced samples/s of the hybrid (regrouping) gfx950 backend
against the predicated compact backend on the smallpt 9-sphere scene,
1024x768 at 64 spp, one thread per sample.

A step renders one whole frame (50,331,648 sample threads) with inputs
resident in HBM.  The JSON line reports the hybrid throughput as `value`,
the compact throughput and active-lane fractions beside it, the roofline of
the dominant kernel (hybrid pass kernel, HIP events around each launch) and
the CPU baseline (the oracle's block interpreter of the same IR kernel, the
reference's scalar VM backend restated, on the host threads of one GPU, a
bounded sample of the same frame; the direct C++ path tracer beside it).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Every execution context of the library has a kernel stream and a copy
# stream, and HIP spreads streams over GPU_MAX_HW_QUEUES hardware queues (4
# by default), sharing queues once there are more streams.  With 4, the
# serving leg run after the batch leg served 2,420 of 3,000 requests/s at a
# p50 of 50 ms (GPU busy 41%): its copies no longer overlapped its kernels.
# With 16 it serves 3,078 requests/s at a p50 of 2.1 ms.  Set before HIP
# initialises.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = ("path-trace samples/sec + active-lane % (hybrid vs compact GPU); "
          "p99 req latency")
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 vector, published (counts packed FP32)
# one non-packed FP32 fma per lane and cycle: 256 CUs x 4 SIMDs x 16 lanes x
# 2 flops x 2.4 GHz (half the published figure; packed FP32 is not used)
FP32_NONPACKED_TFLOPS = 256 * 4 * 16 * 2 * 2.4e9 / 1e12
HBM_PEAK_GBPS = 8000.0    # MI355X HBM3E, published
CPU_THREADS = 16          # host threads of one GPU's share of the test box
POLICY_NAMES = {0: "successor-major", 1: "topo", 2: "sweep"}
W, H, SPP = 1024, 768, 64
SERVE_QPS = 1000.0
SERVE_QPS_HIGH = 3000.0
SERVE_REQUESTS = 1000
SERVE_BUDGET = 4 << 30
SERVE_STEP = 4 << 20


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-compact", action="store_true", help="skip the compact comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-job mixed batch")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-regroup", action="store_true", help="skip the queue-regrouping form")
    ap.add_argument("--no-int", action="store_true", help="skip the fib and Collatz configs")
    ap.add_argument("--no-ref-pt2", action="store_true",
                    help="skip the reference's pt2_const path tracer (1024x1024) in every "
                         "form (built by oracle/build_ref.py into "
                         "oracle/_ref/libloom_refkernels.so)")
    ap.add_argument("--qps", type=float, default=SERVE_QPS, help="serving arrival rate")
    ap.add_argument("--requests", type=int, default=SERVE_REQUESTS)
    return ap.parse_args()


def mixed_batch(steps):
    """The 64-job mixed batch in one persistent launch replayed from a HIP
    graph, against one launch per job; outputs compared bit for bit."""
    import loom_amd
    from loom_amd import abi, autotune, mixed
    mb = mixed.MixedBatch()
    bp = autotune.best_params("batch:" + mixed.SET)
    params = loom_amd.make_params(mode=abi.MODE_HYBRID, stats=False, **bp)
    g = mb.batch().graph(params)
    g.launch()  # warm
    t = time.perf_counter()
    launch_ms = 0.0
    for _ in range(steps):
        launch_ms += g.launch().ms_main
    batch_ms = (time.perf_counter() - t) * 1e3 / steps
    launch_ms /= steps
    got = mb.outputs()
    sep_ms = mb.run_separately()
    want = mb.outputs()
    same = all(np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                              y.view(np.uint32) if y.dtype == np.float32 else y)
               for gj, wj in zip(got, want) for x, y in zip(gj, wj))
    g.destroy()
    kinds = {}
    for name, spec in mb.specs:
        kinds[name] = kinds.get(name, 0) + spec["n"]
    return {"workload": "64 seeded jobs: frames w,h in [64,1024], spp in [1,64]; fib and "
                        "Collatz of 2^12-2^22 threads",
            "threads": mb.threads, "threads_by_kernel": kinds,
            "ms_per_step": batch_ms, "graph_launch_ms": launch_ms,
            "threads_per_s": mb.threads / (batch_ms / 1e3),
            "separate_launches_ms": sep_ms, "speedup_vs_separate": sep_ms / batch_ms,
            "outputs_identical_to_separate": same, "params": bp}


SERVE_CHECKED = 24  # requests per serving run whose outputs are compared
# dispatches between two workgroup sorts on the reference path tracer, for
# 16- and 8-wave workgroups (the fastest of 32-1024 for each size)
WGR_PHASE = {16: 128, 8: 256}


def serving(qps, count):
    """A Poisson stream of mixed requests served with continuous batching.
    Most requests write into shared pinned sinks; a seeded sample of
    SERVE_CHECKED requests gets output columns of its own, compared after
    the run with the same launch run alone (every request of a smaller trace
    is compared in the GPU tests)."""
    import loom_amd
    from loom_amd import abi, serve
    pools = serve.Pools.make()
    specs = serve.poisson_trace(qps=qps, count=count)
    sinks = {"smallpt9_f32": [serve.Pinned(serve.PT_MAX, np.float32) for _ in range(3)],
             "fib_u64": [serve.Pinned(1 << serve.INT_MAX_LOG2, np.uint64)],
             "collatz_u64": [serve.Pinned(1 << serve.INT_MAX_LOG2, np.uint32)]}
    outs = [[p.array[:s.n] for p in sinks[s.kernel]] for s in specs]
    checked = sorted(np.random.default_rng(int(qps)).choice(
        np.arange(8, len(specs)), min(SERVE_CHECKED, len(specs) - 8), replace=False).tolist())
    own = {}
    for i in checked:
        k = loom_amd.kernel(specs[i].kernel)
        own[i] = [serve.Pinned(specs[i].n, dt) for dt in k.outputs]
        outs[i] = [p.array for p in own[i]]
    srv = serve.Server(hbm_budget=SERVE_BUDGET, max_resident=32, step_threads=SERVE_STEP)
    # warm the path (first launches, page-in of the pools)
    srv.run(specs[:8], pools, realtime=False, outputs=outs[:8])
    reqs, st, _ = srv.run(specs, pools, realtime=True, outputs=outs)
    srv.close()
    same = all(q.status == 0 for q in reqs)
    for i in checked:
        alone = loom_amd.run(specs[i].kernel, abi.MODE_HYBRID,
                             pools.inputs_of(specs[i], full=True)).outputs
        same = same and all(np.array_equal(g.view(np.uint8), a.view(np.uint8))
                            for g, a in zip(outs[i], alone))
    summ = serve.latency_summary(reqs, st)
    summ.update({"qps": qps, "hbm_budget": SERVE_BUDGET, "step_threads": SERVE_STEP,
                 "outputs_checked": len(checked), "outputs_identical_to_alone": same,
                 "workload": "Poisson arrivals; half frames 128^2-512^2 at 4-16 spp, half "
                             "fib/Collatz of 2^14-2^20 threads"})
    return summ


def ref_pt2_leg():
    """The reference's own path tracer, pt2_const (blur 5, threshold 0.0005,
    depth 20) on a 1024x1024 frame, compiled unmodified from its checkout by
    this backend (oracle/build_ref.py): frame time, rays/s and active lanes
    of the compact sweep form, the hybrid form (refill, tuned live
    threshold), the hybrid form regrouping the threads of each workgroup by
    block through LDS every WGR_PHASE dispatches (16-wave workgroups, one per
    CU, and 8-wave ones, two per CU, each at its phase), the compact form
    with its register file in memory, and the queue-regrouping form under the
    reference's successor-major policy at a block-population threshold of 32
    (regrouping passes: a wave stops when its best block has no more lanes
    than the threshold and its threads are queued by label for the next
    pass)."""
    import ctypes as C
    import loom_amd
    from loom_amd import abi
    lib = ROOT / "oracle" / "_ref" / "libloom_refkernels.so"
    if not lib.exists():
        return None
    C.CDLL(str(lib))
    k = loom_amd.kernel("ref_pt2_const_b5_t0.0005_d20_1024")
    # ppp = x << 16 | y over the frame (reference smallpt/vm/enc.hpp:30-34)
    ppp = np.array([(x << 16) | y for x in range(1024) for y in range(1024)], np.uint32)
    dev = [loom_amd.DeviceArray.from_numpy(ppp)]
    forms = {
        "compact_sweep": (abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
        "hybrid_live56": (abi.MODE_HYBRID, dict(live_threshold=56)),
        "hybrid_workgroup_regroup": (abi.MODE_HYBRID, dict(
            regroup=abi.REGROUP_WORKGROUP, reconv=abi.RECONV_SWEEP,
            regroup_phase=WGR_PHASE[16])),
        "hybrid_workgroup8_regroup": (abi.MODE_HYBRID, dict(
            regroup=abi.REGROUP_WORKGROUP, reconv=abi.RECONV_SWEEP, regroup_phase=WGR_PHASE[8],
            regroup_waves=8)),
        "compact_memory_registers": (abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP,
                                                            register_file=abi.REGFILE_MEMORY)),
        "regroup_successor_major_thr32": (abi.MODE_HYBRID, dict(
            reconv=abi.RECONV_SUCCESSOR_MAJOR, regroup=True, threshold=32, live_threshold=0)),
    }
    out, first = {}, None
    for name, (mode, params) in forms.items():
        outs = [loom_amd.DeviceArray(len(ppp), dt) for dt in k.outputs]
        r = loom_amd.run(k, mode, dev, outs, params=loom_amd.make_params(mode=mode, **params))
        got = [o.numpy() for o in outs]
        if first is None:
            first = got
        rays = int(got[3].astype(np.uint64).sum())
        out[name] = {"ms": r.ms_total, "rays_per_s": rays / (r.ms_total / 1e3),
                     "active_lane_fraction": r.active_lane_fraction, "passes": r.passes,
                     "outputs_identical": all(np.array_equal(x.view(np.uint32), y.view(np.uint32))
                                              for x, y in zip(got, first))}
        print(f"ref pt2_const 1024x1024 {name}: {r.ms_total:.1f} ms", file=sys.stderr, flush=True)
    return out


def regroup_leg(k, ins, outs, h_img, hyb):
    """The queue-regrouping form on the same frame: every pass spills the
    survivors of each wave to the queue of their next label (one atomic per
    wave and label, rows ranked with mbcnt) and the next pass runs 64-row
    batches of one label.  `live32` and `live48` regroup without in-wave
    refill (a wave stops at that many live lanes and spills them);
    `with_refill` refills free lanes while input lasts, with the tuned live
    threshold and grid, and regroups the survivors once the input is
    exhausted.  Reports the regroup traffic (spilled and filled row bytes)
    over the pass kernels' time."""
    import loom_amd
    from loom_amd import abi
    out = {}
    forms = {f"live{live}": dict(live_threshold=live, refill=False) for live in (32, 48)}
    forms["with_refill"] = dict(live_threshold=hyb["live_threshold"],
                                waves_per_cu=hyb["waves_per_cu"], refill=True)
    for name, kw in forms.items():
        p = loom_amd.make_params(mode=abi.MODE_HYBRID, regroup=True, **kw)
        loom_amd.run(k, abi.MODE_HYBRID, ins, outs, params=p)
        r = loom_amd.run(k, abi.MODE_HYBRID, ins, outs, params=p)
        same = all(np.array_equal(o.numpy().view(np.uint32), x.view(np.uint32))
                   for o, x in zip(outs, h_img))
        moved = r.counters[abi.C_SPILL_BYTES] + r.counters[abi.C_FILL_BYTES]
        gbs = moved / (r.ms_main / 1e3) / 1e9
        out[name] = {
            "ms_total": r.ms_total, "ms_passes": r.ms_main, "passes": r.passes,
            "samples_per_s": W * H * SPP / (r.ms_total / 1e3),
            "active_lane_fraction": r.active_lane_fraction,
            "spill_rows": r.counters[abi.C_SPILL_ROWS], "fill_rows": r.counters[abi.C_FILL_ROWS],
            "regroup_bytes": moved, "regroup_GBps_over_pass_time": gbs,
            "frac_of_hbm_peak": gbs / HBM_PEAK_GBPS, "outputs_identical": same}
    return out


def traffic_from_profile():
    """HBM bytes per launch of the hybrid pass kernel measured with rocprofv3
    (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, scaled by a copy kernel) and
    committed as a short summary; None if it is absent."""
    p = ROOT / "profiles" / "smallpt9_hybrid_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def cpu_baseline():
    """The reference's CPU VM path as restated by the oracle: its block
    interpreter (the scalar backend's semantics, src/vm/bse.hpp:139-319)
    running the same IR kernel, on the GPU's share of host threads, over
    the first 32 rows of the frame (2,097,152 samples, about 10 s of CPU
    work); beside it the oracle's direct C++ path tracer (the same
    per-sample algorithm written directly in C++, no VM), and one-thread
    rates of both."""
    from tests import oracle_lib
    from loom_amd import workloads
    job_word = workloads.pt_job(W, H, SPP)
    n = 32 * W * SPP
    sid = np.arange(n, dtype=np.uint32)
    job = np.full(n, job_word, np.uint32)
    threads = CPU_THREADS
    t = time.perf_counter()
    oracle_lib.run("smallpt9_f32", [sid, job], threads=threads)
    dt = time.perf_counter() - t
    n1 = n // 64
    t = time.perf_counter()
    oracle_lib.run("smallpt9_f32", [sid[:n1], job[:n1]], threads=1)
    dt1 = time.perf_counter() - t
    # the direct tracer: the first half of the frame
    nd = W * H * SPP // 2
    sidd = np.arange(nd, dtype=np.uint32)
    jobd = np.full(nd, job_word, np.uint32)
    t = time.perf_counter()
    oracle_lib.smallpt9_direct(sidd, jobd, threads=threads)
    ddt = time.perf_counter() - t
    nd1 = nd // 16
    t = time.perf_counter()
    oracle_lib.smallpt9_direct(sidd[:nd1], jobd[:nd1], threads=1)
    ddt1 = time.perf_counter() - t
    return {"value": n / dt, "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"first {n} samples (rows 0-31) of the {W}x{H}x{SPP} frame, the "
                      f"oracle's block interpreter of the same IR kernel (the reference's "
                      f"scalar VM backend restated), {threads} threads, {dt:.2f} s wall",
            "single_core": {"value": n1 / dt1, "unit": "samples/s",
                            "sample": f"first {n1} samples, 1 thread, {dt1:.2f} s wall"},
            "algorithm_only": {
                "value": nd / ddt, "unit": "samples/s", "cores": threads,
                "sample": f"first {nd} samples (rows 0-{H // 2 - 1}), the path tracer written "
                          f"directly in C++ (no VM), {threads} threads, {ddt:.2f} s wall",
                "single_core": nd1 / ddt1},
            "cpu_model": cpu_model(), "host_threads_seen": os.cpu_count()}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def integer_configs(steps: int):
    """fib u64 over 2^20 threads and Collatz u64 over 2^24 threads, inputs in
    HBM: the form the autotuner chose, the tuned hybrid form and the compact
    form under the reference policy and the sweep policy, threads/s (timed
    without statistics) and active lanes (from a run with them), outputs
    compared across forms."""
    import loom_amd
    from loom_amd import abi, autotune, workloads
    out = {}
    import hashlib
    fixtures = {"fib_u64": "fib_u64_c1.json", "collatz_u64": "collatz_u64_c2.json"}
    for key, name, cols in (("fib_u64_2^20", "fib_u64", [workloads.fib_inputs(1 << 20)]),
                            ("collatz_u64_2^24", "collatz_u64",
                             [workloads.collatz_inputs(1 << 24)])):
        k = loom_amd.kernel(name)
        n = len(cols[0])
        din = [loom_amd.DeviceArray.from_numpy(c) for c in cols]
        tmode, tparams = autotune.best_form(name)
        forms = {"tuned": autotune.params_of_form(tmode, tparams),
                 "hybrid": (abi.MODE_HYBRID, loom_amd.make_params(
                     mode=abi.MODE_HYBRID, **autotune.best_params(name))),
                 "compact_successor_major": (abi.MODE_COMPACT, loom_amd.make_params(
                     mode=abi.MODE_COMPACT)),
                 "compact_sweep": (abi.MODE_COMPACT, loom_amd.make_params(
                     mode=abi.MODE_COMPACT, reconv=abi.RECONV_SWEEP))}
        entry, first = {"threads": n, "tuned_form": {"mode": tmode, "params": tparams}}, None
        for form, (mode, params) in forms.items():
            dout = [loom_amd.DeviceArray(n, dt) for dt in k.outputs]
            loom_amd.run(k, mode, din, dout, params=params)  # with statistics
            r = loom_amd.run(k, mode, din, dout, params=params)
            params.collect_stats = 0  # timed without
            ms = []
            for _ in range(max(steps, 1)):
                ms.append(loom_amd.run(k, mode, din, dout, params=params).ms_total)
            got = [d.numpy() for d in dout]
            if first is None:
                first = got
            entry[form] = {"ms": float(np.median(ms)), "threads_per_s": n / (np.median(ms) / 1e3),
                           "active_lane_fraction": r.active_lane_fraction,
                           "outputs_identical": all(np.array_equal(x, y)
                                                    for x, y in zip(got, first))}
        # the whole workload's outputs against the committed fixture
        # (computed from the recurrence / brute force, not the oracle)
        fx = ROOT / "tests" / "golden" / fixtures[name]
        if fx.exists():
            f = json.loads(fx.read_text())
            if f["threads"] == n:
                entry["outputs_match_fixture"] = hashlib.sha256(
                    np.ascontiguousarray(first[0]).tobytes()).hexdigest() == f["outputs_sha256"]
        out[key] = entry
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the path "
              f"would shard by sample range or whole jobs with one process per GPU and no "
              f"exchange, and only the one-GPU share is built here", file=sys.stderr)
        return 2
    import loom_amd
    from loom_amd import abi, workloads, autotune

    loom_amd.init(0)
    k = loom_amd.kernel("smallpt9_f32")
    n = W * H * SPP
    ins = [loom_amd.DeviceArray.from_numpy(x) for x in workloads.pt_inputs(W, H, SPP)]
    outs = [loom_amd.DeviceArray(n, np.float32) for _ in range(3)]

    # timed steps collect no statistics (the production setting); the work
    # counters (active lanes, flops) come from one more run of each form
    # with statistics, whose outputs are the same
    hyb = autotune.best_params("smallpt9_f32")
    hp = loom_amd.make_params(mode=abi.MODE_HYBRID, stats=False, **hyb)
    hp_stats = loom_amd.make_params(mode=abi.MODE_HYBRID, **hyb)
    cp = loom_amd.make_params(mode=abi.MODE_COMPACT, stats=False)

    def steps(mode, params, count):
        last = None
        t0 = time.perf_counter()
        for _ in range(count):
            last = loom_amd.run(k, mode, ins, outs, params=params)
        abi.check(abi.load().loom_synchronize(), "sync")
        return (time.perf_counter() - t0) * 1e3 / max(count, 1), last

    steps(abi.MODE_HYBRID, hp, a.warmup)
    h_ms, h_last = steps(abi.MODE_HYBRID, hp, a.steps)
    h_img = [o.numpy() for o in outs]
    h_stats = loom_amd.run(k, abi.MODE_HYBRID, ins, outs, params=hp_stats)

    result = {
        "metric": METRIC,
        "value": n / (h_ms / 1e3),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": h_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"smallpt9 path tracer {W}x{H} at {SPP} spp, one thread per "
                               f"sample ({n} threads), 9-sphere scene, max depth 32",
                   "width": W, "height": H, "spp": SPP, "samples": n,
                   "backend": "hybrid", "hybrid_params": hyb},
        "active_lane_fraction": h_stats.active_lane_fraction,
    }
    # the measured frame against the committed whole-frame fixture (SHA-256
    # of each radiance channel, from the direct CPU path tracer)
    fixture = ROOT / "tests" / "golden" / "smallpt9_1024x768x64.json"
    if fixture.exists():
        import hashlib
        f = json.loads(fixture.read_text())
        if (f["w"], f["h"], f["spp"]) == (W, H, SPP):
            result["outputs_match_fixture"] = all(
                hashlib.sha256(np.ascontiguousarray(x.view(np.uint32)).tobytes()).hexdigest() == d
                for x, d in zip(h_img, f["outputs_sha256"]))

    if not a.no_compact:
        same = True
        for key, params, label in (
                ("compact", cp, "successor-major (reference policy)"),
                ("compact_sweep", loom_amd.make_params(mode=abi.MODE_COMPACT, stats=False,
                                                       reconv=abi.RECONV_SWEEP),
                 "sweep (every block in topological order)")):
            steps(abi.MODE_COMPACT, params, 1)
            c_ms, _ = steps(abi.MODE_COMPACT, params, a.steps)
            c_img = [o.numpy() for o in outs]
            params.collect_stats = 1
            c_last = loom_amd.run(k, abi.MODE_COMPACT, ins, outs, params=params)
            same = same and all(np.array_equal(x.view(np.uint32), y.view(np.uint32))
                                for x, y in zip(h_img, c_img))
            result[key] = {"value": n / (c_ms / 1e3), "ms_per_step": c_ms,
                           "active_lane_fraction": c_last.active_lane_fraction,
                           "reconvergence": label}
            result["hybrid_over_" + key] = c_ms / h_ms
        result["outputs_identical"] = same

    # roofline of the dominant kernel: the hybrid pass kernel
    launches = max(h_last.main_launches, 1)
    launch_ms = h_last.ms_main / launches
    flops = h_stats.counters[abi.C_ACTIVE_FLOPS] / max(h_stats.main_launches, 1)
    achieved = flops / (launch_ms / 1e3) / 1e12
    result["roofline"] = {
        "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic_from_profile(),
        "kernel": f"hybrid_pass_kernel<smallpt9_f32, {POLICY_NAMES[hyb['reconv']]}>",
        "launch_ms": launch_ms, "launches_per_step": launches,
        "algorithmic_flops_per_launch": flops,
        "peak_nonpacked": FP32_NONPACKED_TFLOPS,
        "frac_of_nonpacked": achieved / FP32_NONPACKED_TFLOPS,
        "note": "vector-ALU bound, no MFMA in this path: useful FP32 flops of active lanes "
                "(fma = 2, device counter) per launch over the launch's HIP-event time; peak "
                "is the published FP32 vector figure, which counts packed FP32; packed FP32 "
                "is not used, so peak_nonpacked (one fma per lane and cycle) is the "
                "attainable ceiling"}

    if not a.no_regroup:
        rg = regroup_leg(k, ins, outs, h_img, hyb)
        for key in ("compact", "compact_sweep"):
            if key in result:
                rg["with_refill"][f"over_{key}"] = (result[key]["ms_per_step"]
                                                    / rg["with_refill"]["ms_total"])
        result["regroup"] = rg
    if not a.no_int:
        result["integer_configs"] = integer_configs(a.steps)
    if not a.no_ref_pt2:
        try:
            result["ref_pt2_const_1024"] = ref_pt2_leg()
        except Exception as e:  # the reference's kernels are optional here
            result["ref_pt2_const_1024"] = {"error": str(e)}
    if not a.no_batch:
        result["batch"] = mixed_batch(max(1, min(a.steps, 3)))
    if not a.no_serve:
        sv = serving(a.qps, a.requests)
        result["serving"] = sv
        result["p50_latency_ms"] = sv["p50_ms"]
        result["p99_latency_ms"] = sv["p99_ms"]
        if a.qps == SERVE_QPS:
            # a second load point, near saturation
            hi = serving(SERVE_QPS_HIGH, a.requests)
            result["serving_high_load"] = {k: hi[k] for k in (
                "qps", "requests_per_s", "p50_ms", "p99_ms", "max_ms", "gpu_busy",
                "threads_per_s", "peak_resident", "outputs_checked",
                "outputs_identical_to_alone")}

    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
