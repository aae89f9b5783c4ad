"""The reference's own path tracers (pt_const, pt2_const at two of its
benchmarked configurations) on the GPU through the C ABI, every form,
bit-exact against the oracle interpreter running the same kernels; frame
times printed beside the reference's published single-core CPU times."""
import time

import numpy as np
import pytest

import loom_amd
from loom_amd import abi
from tests import oracle_lib as O
from tests.test_refkernels import REF_LIB, frame_inputs, load_ref_kernels

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (REF_LIB.exists() and O.REF_PATH.exists()),
                                 reason="reference kernels not built (no reference checkout)")]

FORMS = [
    ("compact_sm", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
    ("compact_memory", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP,
                                              register_file=abi.REGFILE_MEMORY)),
    ("hybrid_sweep", abi.MODE_HYBRID, {}),
    ("hybrid_sm_regroup", abi.MODE_HYBRID, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR,
                                                regroup=True, live_threshold=32)),
    ("hybrid_workgroup_regroup", abi.MODE_HYBRID, dict(regroup=abi.REGROUP_WORKGROUP,
                                                       reconv=abi.RECONV_SWEEP,
                                                       regroup_phase=128)),
]
# published single-core Zen 4 frame times, s: predicated (compact), hybrid
# (reference site-docs/benchmarks/results/a0d934323c/5c42ee70d3/*.csv)
PUBLISHED = {"ref_pt_const_b1_t0.0001_d2": (1.207, 0.784),
             "ref_pt2_const_b5_t0.0005_d20": (2.169, 1.748)}


@pytest.mark.parametrize("name", list(PUBLISHED))
def test_reference_kernel_parity(gpu, name):
    load_ref_kernels()
    ppp = frame_inputs()
    print(f"[gpu] reference kernel {name} 40x40", flush=True)
    want, err = O.run(name, [ppp], threads=16, path=O.REF_PATH)
    assert err == 0
    k = loom_amd.kernel(name)
    dev = [loom_amd.DeviceArray.from_numpy(ppp)]
    for fname, mode, params in FORMS:
        print(f"[gpu] {name} {fname}", flush=True)
        p = loom_amd.make_params(mode=mode, **params)
        outs = [loom_amd.DeviceArray(len(ppp), dt) for dt in k.outputs]
        r = loom_amd.run(k, mode, dev, outs, params=p)
        assert r.errors == 0
        t0 = time.perf_counter()
        r = loom_amd.run(k, mode, dev, outs, params=p)
        wall = time.perf_counter() - t0
        for g, w_ in zip(outs, want):
            assert np.array_equal(g.numpy().view(np.uint32), w_.view(np.uint32)), (name, fname)
        print(f"[gpu] {name} {fname}: {r.ms_total:.2f} ms per frame (wall {wall * 1e3:.2f} ms), "
              f"active lanes {r.active_lane_fraction:.3f}; published CPU core "
              f"{PUBLISHED[name][0]:.3f} s compact, {PUBLISHED[name][1]:.3f} s hybrid", flush=True)


def test_reference_kernel_gpu_sized_frame(gpu):
    """pt2_const (blur 5, threshold 0.0005, depth 20) on a 1024x1024 frame:
    a spread of pixels bit-exact against the oracle in the compact and the
    hybrid form, then the whole frame timed, in rays per second beside the
    reference's published single-core rate for the same kernel (its 40x40
    frame's rays, counted by the oracle, over the published frame time)."""
    load_ref_kernels()
    name = "ref_pt2_const_b5_t0.0005_d20_1024"
    k = loom_amd.kernel(name)
    full = frame_inputs(1024, 1024)
    pick = full[np.random.default_rng(7).choice(len(full), 384, replace=False)]
    print(f"[gpu] {name}: oracle on {len(pick)} pixels", flush=True)
    want, err = O.run(name, [pick], threads=16, path=O.REF_PATH)
    assert err == 0
    forms = [("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
             ("compact_memory", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP,
                                                       register_file=abi.REGFILE_MEMORY)),
             ("hybrid_sweep", abi.MODE_HYBRID, {}),
             ("hybrid_workgroup_regroup", abi.MODE_HYBRID,
              dict(regroup=abi.REGROUP_WORKGROUP, reconv=abi.RECONV_SWEEP, regroup_phase=128)),
             ("hybrid_workgroup8_regroup", abi.MODE_HYBRID,
              dict(regroup=abi.REGROUP_WORKGROUP, reconv=abi.RECONV_SWEEP, regroup_phase=256,
                   regroup_waves=8))]
    for fname, mode, params in forms:
        got = loom_amd.run(k, mode, [pick], params=loom_amd.make_params(mode=mode, **params))
        assert got.errors == 0
        for g, w_ in zip(got.outputs, want):
            assert np.array_equal(g.view(np.uint32), w_.view(np.uint32)), fname
    # published rate: rays of the 40x40 frame over its published time
    small = "ref_pt2_const_b5_t0.0005_d20"
    (_, _, _, rc40, _), err = O.run(small, [frame_inputs()], threads=16, path=O.REF_PATH)
    assert err == 0
    cpu_rays = int(rc40.astype(np.uint64).sum())
    dev = [loom_amd.DeviceArray.from_numpy(full)]
    for fname, mode, params in forms:
        p = loom_amd.make_params(mode=mode, **params)
        outs = [loom_amd.DeviceArray(len(full), dt) for dt in k.outputs]
        loom_amd.run(k, mode, dev, outs, params=p)
        r = loom_amd.run(k, mode, dev, outs, params=p)
        rays = int(outs[3].numpy().astype(np.uint64).sum())
        print(f"[gpu] {name} {fname}: {r.ms_total:.1f} ms per 1024x1024 frame, "
              f"{rays / (r.ms_total / 1e3) / 1e9:.3f} G rays/s, active lanes "
              f"{r.active_lane_fraction:.3f}; reference single core (Zen 4): "
              f"{cpu_rays / PUBLISHED[small][0] / 1e6:.3f} M rays/s compact, "
              f"{cpu_rays / PUBLISHED[small][1] / 1e6:.3f} M rays/s hybrid", flush=True)
