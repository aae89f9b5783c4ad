"""The GPU backend against the committed golden fixtures at full size: the
fib workload (2^20 threads), the Collatz workload (2^24 threads) and a
32x24x8 path-traced frame, in the compact and hybrid forms."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import loom_amd
from loom_amd import abi, workloads

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"
FORMS = [(abi.MODE_COMPACT, {}), (abi.MODE_HYBRID, {}),
         (abi.MODE_HYBRID, dict(live_threshold=32, regroup=True))]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(name, ins, mode, params):
    k = loom_amd.kernel(name)
    d = [loom_amd.DeviceArray.from_numpy(a) for a in ins]
    o = [loom_amd.DeviceArray(len(ins[0]), dt) for dt in k.outputs]
    print(f"[gpu] golden {name} n={len(ins[0])} mode={mode} {params}", flush=True)
    r = loom_amd.run(k, mode, d, o, params=loom_amd.make_params(mode=mode, **params))
    assert r.errors == 0
    return [x.numpy() for x in o]


@pytest.mark.parametrize("mode,params", FORMS)
def test_fib_c1(gpu, mode, params):
    f = json.loads((G / "fib_u64_c1.json").read_text())
    n = workloads.fib_inputs(f["threads"])
    (c,) = run("fib_u64", [n], mode, params)
    assert sha(c) == f["outputs_sha256"]


@pytest.mark.parametrize("mode,params", FORMS)
def test_collatz_c2(gpu, mode, params):
    f = json.loads((G / "collatz_u64_c2.json").read_text())
    x = workloads.collatz_inputs(f["threads"])
    (s,) = run("collatz_u64", [x], mode, params)
    assert sha(s) == f["outputs_sha256"]
    assert int(s.max()) == f["max_steps"]


@pytest.mark.parametrize("mode,params", FORMS)
def test_smallpt9_frame(gpu, mode, params):
    z = np.load(G / "smallpt9_32x24x8.npz")
    ins = workloads.pt_inputs(int(z["w"]), int(z["h"]), int(z["spp"]))
    out = run("smallpt9_f32", ins, mode, params)
    for got, key in zip(out, ("lx", "ly", "lz")):
        assert np.array_equal(got.view(np.uint32), z[key])


FRAME_FORMS = [("compact_sm", abi.MODE_COMPACT, {}),
               ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
               ("hybrid_tuned", abi.MODE_HYBRID, None)]


@pytest.mark.parametrize("fname,mode,params", FRAME_FORMS)
def test_smallpt9_full_frame(gpu, fname, mode, params):
    """The whole path-tracing workload frame (1024x768 at 64 spp, 50,331,648
    sample threads) bit for bit against the fixture from the direct C++
    restatement, per radiance channel."""
    from loom_amd import autotune
    f = json.loads((G / "smallpt9_1024x768x64.json").read_text())
    ins = workloads.pt_inputs(f["w"], f["h"], f["spp"])
    if params is None:
        params = autotune.best_params("smallpt9_f32")
    out = run("smallpt9_f32", ins, mode, params)
    print(f"[gpu] full frame {fname}: comparing", flush=True)
    for c, got in enumerate(out):
        assert sha(got.view(np.uint32)) == f["outputs_sha256"][c], (fname, c)
