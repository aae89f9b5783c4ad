"""The sin/cos polynomials of the op contract (the reference SIMD backends',
src/vm/simd.hpp:955-1100) against vectors evaluated in exact arithmetic from
the reference's coefficient tables (tests/golden/sincos.json, written by
tests/golden/make_sincos.py): the oracle on the CPU, the product on the GPU,
bit for bit; and the coefficient tables of both restatements against the
reference's text where the checkout exists."""
from __future__ import annotations

import json
import re
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from tests import literal_ref as R
from tests import oracle_lib as O
from tests.golden import make_sincos as M

ROOT = Path(__file__).resolve().parent.parent
FIX = json.loads((ROOT / "tests" / "golden" / "sincos.json").read_text())
KINDS = {"f32": (np.float32, np.uint32), "f64": (np.float64, np.uint64)}
SIMD_HPP = Path("/root/reference/src/vm/simd.hpp")


def vectors(kind):
    dt, ut = KINDS[kind]
    row = FIX[kind]
    xs = np.array(row["x_bits"], dtype=ut).view(dt)
    return xs, np.array(row["sin_bits"], dtype=ut), np.array(row["cos_bits"], dtype=ut)


def tables_in_source(path: Path, fn_names: dict) -> dict:
    """Coefficient values (k1 first) of a restatement's sin/cos, read from
    its source: every decimal literal of the function's float / double
    branch, in k order (the device code lists them highest first)."""
    text = path.read_text()
    out = {}
    num = r"-?\d+\.\d*(?:e-?\d+)?"
    for (fn, kind), (marker, reverse) in fn_names.items():
        start = text.index(marker)
        end = text.index("\n}\n", start)
        body = text[start:end]
        if "if constexpr" in body:
            flt, _, dbl = body.partition("else")
            body = flt if kind == "f32" else dbl
        vals = re.findall(num + ("f" if kind == "f32" else r"(?![\df])"), body)
        vals = [v.rstrip("f") for v in vals]
        if reverse:
            # fma_(x2, k_n, k_{n-1}), then fma_(p, x2, k_{n-2}) ... k1
            vals = vals[::-1]
        out[(fn, kind)] = [R.round_to(Fraction(v), kind) for v in vals]
    return out


DEVICE = tables_in_source(ROOT / "loom_amd" / "csrc" / "device" / "ops.hpp", {
    ("sin", "f32"): ("poly_sin(T x)", True), ("sin", "f64"): ("poly_sin(T x)", True),
    ("cos", "f32"): ("poly_cos(T x)", True), ("cos", "f64"): ("poly_cos(T x)", True)})
ORACLE = tables_in_source(ROOT / "oracle" / "ops.hpp", {
    ("sin", "f32"): ("sin_poly(float x)", False), ("sin", "f64"): ("sin_poly(double x)", False),
    ("cos", "f32"): ("cos_poly(float x)", False), ("cos", "f64"): ("cos_poly(double x)", False)})


@pytest.mark.parametrize("kind", ["f32", "f64"])
def test_oracle_matches_vectors(kind):
    L = O.lib()
    dt, ut = KINDS[kind]
    xs, s_bits, c_bits = vectors(kind)
    for fn, want in (("sin", s_bits), ("cos", c_bits)):
        f = getattr(L, f"oracle_{fn}_{kind}")
        got = np.array([f(float(x)) for x in xs], dtype=dt).view(ut)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (fn, kind, xs[bad[:4]], got[bad[:4]], want[bad[:4]])


def test_tables_have_reference_lengths():
    for t in (DEVICE, ORACLE):
        assert [len(t[(f, k)]) for f in ("sin", "cos") for k in ("f32", "f64")] == [8, 17, 9, 17]


def test_device_and_oracle_tables_agree():
    assert DEVICE == ORACLE


def test_tables_match_reference_text():
    """Digit for digit (as values of the type) against the reference's
    DECLK lines, where the checkout exists."""
    if not SIMD_HPP.exists():
        pytest.skip("reference checkout absent")
    ref = M.coefficient_tables(SIMD_HPP)
    assert DEVICE == ref
    assert ORACLE == ref


def test_evaluation_order_reproduces_vectors():
    """The device tables, evaluated in the nested-fma order the product
    uses, reproduce the committed vectors (sampled)."""
    for kind in ("f32", "f64"):
        xs, s_bits, c_bits = vectors(kind)
        for i in range(0, len(xs), 97):
            for fn, want in (("sin", s_bits), ("cos", c_bits)):
                v = M.poly(fn, Fraction(float(xs[i])), DEVICE[(fn, kind)], kind)
                if v != 0:
                    assert R.bits(v, kind) == int(want[i]), (fn, kind, xs[i])


def test_vectors_catch_a_one_ulp_coefficient_slip():
    """A planted mistake in a coefficient (k2 of the float sine moved by one
    ulp) changes outputs the vectors hold, so the GPU and oracle checks
    against them would fail."""
    xs, s_bits, _ = vectors("f32")
    ks = list(DEVICE[("sin", "f32")])
    ks[1] = ks[1] + Fraction(2) ** (-26)  # one ulp of |k2| ~ 0.1667 in binary32
    diff = 0
    for i in range(9, 9 + 256):
        v = M.poly("sin", Fraction(float(xs[i])), ks, "f32")
        diff += R.bits(v, "f32") != int(s_bits[i])
    assert diff > 0


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["compact", "hybrid"])
def test_gpu_matches_vectors(gpu, kind, mode):
    import loom_amd
    from loom_amd import abi
    xs, s_bits, c_bits = vectors(kind)
    _, ut = KINDS[kind]
    m = abi.MODE_COMPACT if mode == "compact" else abi.MODE_HYBRID
    k = loom_amd.kernel(f"sincos_{kind}")
    din = [loom_amd.DeviceArray.from_numpy(xs)]
    dout = [loom_amd.DeviceArray(len(xs), xs.dtype) for _ in range(2)]
    loom_amd.run(k, m, din, dout, params=loom_amd.make_params(mode=m))
    for d, want, fn in zip(dout, (s_bits, c_bits), ("sin", "cos")):
        got = d.numpy().view(ut)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (fn, kind, mode, xs[bad[:4]])
