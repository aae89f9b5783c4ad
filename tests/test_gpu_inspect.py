"""Compiled resources and the inspector's counter run.  Every kernel keeps
its register file in VGPRs: no form or policy uses scratch memory (a
regression guard for code shapes that demote the register file)."""
import pytest

import loom_amd
from loom_amd import inspector as inspect

pytestmark = pytest.mark.gpu


# kernels registered with an occupancy hint (LOOM_HIP_WAVES_HINT) and the
# scratch bytes per lane that the hint may cost them: the path tracer's pass
# kernels are held to 96 VGPRs (five waves per SIMD), 32-40 bytes spilled,
# with and without the regroup paths; the reference's pt2_const (registered
# when its library is loaded) is held to 128 without the regroup paths:
# nothing spilled under the sweep policy, up to 160 bytes under the other
# two; its workgroup-regrouping kernels (128 VGPRs at most) spill 188 bytes
# under successor-major
HINTED_SCRATCH = {"smallpt9_f32": 48, "ref_pt2_const_b5_t0.0005_d20": 192,
                  "ref_pt2_const_b5_t0.0005_d20_1024": 192}


def test_no_kernel_uses_scratch(gpu):
    print("[gpu] kernel resources", flush=True)
    for k in loom_amd.kernels():
        res = inspect.resources(k.name)
        for form, r in res.items():
            assert 0 < r["vgprs"] <= 512, (k.name, form, r)
            assert r["scratch_bytes"] <= HINTED_SCRATCH.get(k.name, 0), (k.name, form, r)
            if k.name in HINTED_SCRATCH and form in ("hybrid/sweep", "hybrid/topo",
                                                     "hybrid/successor-major"):
                assert r["vgprs"] <= 128, (k.name, form, r)


def test_inspector_counter_run(gpu):
    print("[gpu] inspector counters", flush=True)
    c = inspect.counters("collatz_u64", 4096)
    for form in ("compact", "hybrid"):
        assert c[form]["admitted"] in (0, 4096)
        assert 0 < c[form]["active_lane_fraction"] <= 1
        assert c[form]["active_ops"] <= c[form]["slot_ops"]
    # the same threads run the same blocks in both forms
    assert c["compact"]["thread_steps"] == c["hybrid"]["thread_steps"]


def test_pass_kernel_without_regroup_is_not_larger(gpu):
    """Runs that do not regroup launch a pass kernel without the queue paths;
    it never needs more registers than the one that regroups."""
    print("[gpu] pass kernel resources with and without regroup", flush=True)
    for name in ("smallpt9_f32", "collatz_u64", "fp_mix_f32"):
        res = inspect.resources(name)
        for pol in ("successor-major", "topo", "sweep"):
            plain, rg = res[f"hybrid/{pol}"], res[f"hybrid-regroup/{pol}"]
            print(f"[gpu] {name} {pol}: {plain['vgprs']} / {rg['vgprs']} VGPRs", flush=True)
            assert plain["vgprs"] <= rg["vgprs"], (name, pol)
