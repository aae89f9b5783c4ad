"""Square root and reciprocal square root of the op contract (src/vm/ops.hpp
sqrt / rsqrt: std::sqrt and 1 / std::sqrt) over the whole input range, bit
for bit against numpy's correctly rounded sqrt and division: the oracle on
the CPU, the product on the GPU.

The inputs are laid out in groups of 64 consecutive elements, one wave of the
first pass: groups of ordinary positive values only, groups that add one
zero, subnormal, negative, infinite, NaN or very small or very large value,
and groups of random bit patterns, so that both the wave-uniform fast branch
of the device's f32 roots and its general branch produce results."""
from __future__ import annotations

import numpy as np
import pytest

from tests import oracle_lib as O

KINDS = {"f32": (np.float32, np.uint32), "f64": (np.float64, np.uint64)}
WAVE = 64


def special(dt):
    fi = np.finfo(dt)
    tiny = fi.smallest_subnormal
    return np.array([0.0, -0.0, tiny, -tiny, fi.tiny, fi.tiny / 2, fi.max, -fi.max, np.inf,
                     -np.inf, np.nan, -1.0, 2.0 ** -96, np.nextafter(dt(2.0 ** -96), dt(0)),
                     2.0 ** 126, np.nextafter(dt(2.0 ** 126), dt(np.inf)), 2.0 ** 127, 1.0,
                     4.0, 0.25, 2.0 ** -126, 2.0 ** -149 if dt is np.float32 else 2.0 ** -1074],
                    dtype=dt)


def inputs(kind, groups=48, seed=31):
    dt, ut = KINDS[kind]
    rng = np.random.default_rng(seed)
    bits = 32 if dt is np.float32 else 64
    out = []
    sp = special(dt)
    for g in range(groups):
        if g % 3 == 0:
            # positive, normal, exponent spread over the fast range
            e = rng.uniform(-90, 120, WAVE)
            v = (rng.uniform(1, 2, WAVE) * 2.0 ** e).astype(dt)
        elif g % 3 == 1:
            e = rng.uniform(-90, 120, WAVE)
            v = (rng.uniform(1, 2, WAVE) * 2.0 ** e).astype(dt)
            v[rng.integers(WAVE)] = sp[(g // 3) % len(sp)]
        else:
            v = rng.integers(0, 2 ** bits, WAVE, dtype=np.uint64).astype(ut).view(dt)
        out.append(v)
    # every exact square of a small integer and its neighbours
    sq = np.arange(1, 4097, dtype=np.float64) ** 2
    out.append(sq.astype(dt))
    out.append(np.nextafter(sq.astype(dt), dt(np.inf)))
    out.append(np.nextafter(sq.astype(dt), dt(0)))
    x = np.concatenate(out)
    return np.ascontiguousarray(x[: len(x) // WAVE * WAVE])


def expected(x):
    dt = x.dtype.type
    with np.errstate(all="ignore"):
        s = np.sqrt(x)
        r = (dt(1) / np.sqrt(x)).astype(dt)
    return s, r


def same(got, want):
    """Bitwise equality, any NaN equal to any NaN."""
    ut = KINDS["f32" if got.dtype == np.float32 else "f64"][1]
    nan = np.isnan(got) & np.isnan(want)
    return (got.view(ut) == want.view(ut)) | nan


@pytest.mark.parametrize("kind", ["f32", "f64"])
def test_inputs_reach_both_branches(kind):
    x = inputs(kind)
    g = x.reshape(-1, WAVE)
    dt = x.dtype.type
    with np.errstate(invalid="ignore"):
        fast = ((g >= dt(2.0 ** -96)) & (g <= dt(2.0 ** 126))).all(axis=1)
    assert fast.sum() >= 10 and (~fast).sum() >= 10


@pytest.mark.parametrize("kind", ["f32", "f64"])
def test_oracle_roots(kind):
    x = inputs(kind)
    (s, r), _ = O.run(f"roots_{kind}", [x])
    ws, wr = expected(x)
    for got, want, fn in ((s, ws, "sqrt"), (r, wr, "rsqrt")):
        ok = same(got, want)
        assert ok.all(), (fn, x[~ok][:4])


def test_plant_wrong_rounding_fails():
    """A root one ulp off (a truncating square root) fails the comparison."""
    x = inputs("f32")
    ws, _ = expected(x)
    with np.errstate(all="ignore"):
        trunc = np.sqrt(x.astype(np.float64))
        t32 = trunc.astype(np.float32)
        t32 = np.where(t32.astype(np.float64) > trunc, np.nextafter(t32, np.float32(0)), t32)
    assert not same(t32.astype(np.float32), ws).all()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["compact", "hybrid"])
def test_gpu_roots(gpu, kind, mode):
    import loom_amd
    from loom_amd import abi
    x = inputs(kind)
    m = abi.MODE_COMPACT if mode == "compact" else abi.MODE_HYBRID
    k = loom_amd.kernel(f"roots_{kind}")
    din = [loom_amd.DeviceArray.from_numpy(x)]
    dout = [loom_amd.DeviceArray(len(x), x.dtype) for _ in range(2)]
    loom_amd.run(k, m, din, dout, params=loom_amd.make_params(mode=m))
    ws, wr = expected(x)
    for d, want, fn in zip(dout, (ws, wr), ("sqrt", "rsqrt")):
        got = d.numpy()
        ok = same(got, want)
        assert ok.all(), (fn, kind, mode, x[~ok][:4], got[~ok][:4], want[~ok][:4])
