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
