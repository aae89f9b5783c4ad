"""Division by literals.  The backend forms x / d for a literal d = ±2^k
(with a normal reciprocal) as x · (1/d), the same real number rounded once;
the oracle divides.  Both must equal numpy's correctly rounded division bit
for bit, over ordinary values and zeros, subnormals, infinities and NaN."""
from __future__ import annotations

import numpy as np
import pytest

from tests import oracle_lib as O
from tests.test_roots import KINDS, same

DIVISORS = {"q1": 0.5, "q2": 16777216.0, "q3": -0.125, "q4": 3.0, "q5": 0.1}


def inputs(kind, n=1 << 14, seed=41):
    dt, ut = KINDS[kind]
    rng = np.random.default_rng(seed)
    bits = 32 if dt is np.float32 else 64
    fi = np.finfo(dt)
    x = rng.integers(0, 2 ** bits, n, dtype=np.uint64).astype(ut).view(dt)
    x[: n // 2] = (rng.uniform(-1, 1, n // 2) * 2.0 ** rng.uniform(-30, 30, n // 2)).astype(dt)
    sp = np.array([0.0, -0.0, fi.smallest_subnormal, -fi.smallest_subnormal, fi.tiny, fi.max,
                   -fi.max, np.inf, -np.inf, np.nan, 1.0, 3.0], dtype=dt)
    x[: len(sp)] = sp
    return x


def expected(x):
    dt = x.dtype.type
    with np.errstate(all="ignore"):
        out = [(x / dt(DIVISORS[k])).astype(dt) for k in DIVISORS]
        out.append((dt(1) / x).astype(dt))
    return out


@pytest.mark.parametrize("kind", ["f32", "f64"])
def test_oracle_literal_div(kind):
    x = inputs(kind)
    got, _ = O.run(f"literal_div_{kind}", [x])
    want = expected(x)
    for name, g, w in zip(list(DIVISORS) + ["q6"], got, want):
        ok = same(g, w)
        assert ok.all(), (name, x[~ok][:4], g[~ok][:4], w[~ok][:4])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["compact", "hybrid"])
def test_gpu_literal_div(gpu, kind, mode):
    import loom_amd
    from loom_amd import abi
    x = inputs(kind)
    m = abi.MODE_COMPACT if mode == "compact" else abi.MODE_HYBRID
    k = loom_amd.kernel(f"literal_div_{kind}")
    din = [loom_amd.DeviceArray.from_numpy(x)]
    dout = [loom_amd.DeviceArray(len(x), x.dtype) for _ in range(6)]
    loom_amd.run(k, m, din, dout, params=loom_amd.make_params(mode=m))
    want = expected(x)
    for name, d, w in zip(list(DIVISORS) + ["q6"], dout, want):
        g = d.numpy()
        ok = same(g, w)
        assert ok.all(), (name, kind, mode, x[~ok][:4], g[~ok][:4], w[~ok][:4])
