"""Constant selection (ops::sct) with indices inside and outside the constant
list.  sct(i, c0, ..., cN-1) is c_i for i < N and 0 otherwise (reference
operators.hpp `sct`); the backend reads lists of at least four 4-byte
constants from a table in LDS (index clamped to the list's trailing 0) and
shorter or 8-byte lists through compares and selects.  The oracle and every
GPU form must equal the closed form below bit for bit, for indices 0..N-1,
just past the list and at the top of the index type's range."""
from __future__ import annotations

import numpy as np
import pytest

from tests import oracle_lib as O

A = [1.5, -2.25, 3e3, 0.125, -7, 9.5]
B = [7, 0xFFFF, 3, 0x12345678, 1]
C = [0.5, 0.25, -1, 2]
S3 = [4, 5, 6]
KINDS = {"f32": (np.float32, np.uint32), "f64": (np.float64, np.uint64)}


def inputs(kind, n=4099, seed=43):
    _, it = KINDS[kind]
    rng = np.random.default_rng(seed)
    top = np.iinfo(it).max
    x = rng.integers(0, 12, n).astype(it)
    sp = [0, 1, 2, 3, 4, 5, 6, 7, 8, top, top - 1, top - 2, 1 << 31, (1 << 31) - 1]
    if it is np.uint64:
        sp += [1 << 32, (1 << 32) + 1, (1 << 32) + 2, 1 << 63]
    x[: len(sp)] = np.array(sp, dtype=it)
    return x


def pick(table, idx, dt):
    out = np.zeros(len(idx), dtype=dt)
    ok = idx < len(table)
    out[ok] = np.array(table, dtype=dt)[idx[ok].astype(np.int64)]
    return out


def expected(kind, x):
    ft, it = KINDS[kind]
    y = (x - it(2)).astype(it)  # modular, as the IR's unsigned subtraction
    return [pick(A, x, ft), pick(B, x, it), pick(C, y, ft), pick(S3, x, ft)]


def check(got, want, where):
    for name, g, w in zip(("a", "b", "c", "s3"), got, want):
        gb, wb = np.asarray(g).view(f"u{w.itemsize}"), w.view(f"u{w.itemsize}")
        bad = np.flatnonzero(gb != wb)
        assert not len(bad), (where, name, bad[:4], np.asarray(g)[bad[:4]], w[bad[:4]])


@pytest.mark.parametrize("kind", list(KINDS))
def test_oracle_sct_range(kind):
    x = inputs(kind)
    got, err = O.run(f"sct_range_{kind}", [x])
    assert err == 0
    check(got, expected(kind, x), ("oracle", kind))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", list(KINDS))
def test_gpu_sct_range(gpu, kind):
    import loom_amd
    from loom_amd import abi
    x = inputs(kind)
    want = expected(kind, x)
    k = loom_amd.kernel(f"sct_range_{kind}")
    forms = [("compact successor-major", abi.MODE_COMPACT, {}),
             ("compact sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
             ("compact memory", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP,
                                                       register_file=abi.REGFILE_MEMORY)),
             ("hybrid", abi.MODE_HYBRID, {}),
             ("hybrid regroup", abi.MODE_HYBRID, dict(regroup=True, live_threshold=32)),
             ("hybrid workgroup", abi.MODE_HYBRID, dict(regroup=abi.REGROUP_WORKGROUP,
                                                        regroup_waves=8))]
    for name, mode, params in forms:
        print(f"[gpu] sct_range_{kind} {name}", flush=True)
        r = loom_amd.run(k, mode, [x], params=loom_amd.make_params(mode=mode, **params))
        assert r.errors == 0
        check(r.outputs, want, (kind, name))
