"""Writes tests/golden/sincos.json: input/output vectors of the reference
SIMD backends' polynomial sine and cosine (src/vm/simd.hpp:955-1100), which
the op contract adopts for `ops::sin` / `ops::cos`.

    python tests/golden/make_sincos.py [--ref /root/reference]

The coefficients are read from the reference checkout's source text (the
`DECLK(kN, value)` lines of the float and double branches of the sine and
cosine remaps), rounded to the branch's type as its literals are (an `f`
suffix makes a float literal).  The polynomials are evaluated as the
reference's `detail::nest_fma` nests them (simd.hpp:193-201):

    nest(x2, k1, k2)        = fma(x2, k2, k1)
    nest(x2, k1, k2, ks...) = fma(nest(x2, k2, ks...), x2, k1)
    sin(x) = x * nest(x*x, k1..k8)   (float),  k1..k17 (double)
    cos(x) =     nest(x*x, k1..k9)   (float),  k1..k17 (double)

with every fma computed exactly as a rational and rounded once, and every
product rounded once (round-to-nearest-even), independently of the C++ and
HIP restatements under test.  The fixture holds inputs and expected output
bits only.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import re
import struct
import sys
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
from tests import literal_ref as R  # noqa: E402

N_F32 = 4096
N_F64 = 4096
SEED = 0x51C05


def coefficient_tables(simd_hpp: Path) -> dict:
    """{("sin"|"cos", "f32"|"f64"): [k1, k2, ...]} as exact values of the
    type, from the reference's source text."""
    text = simd_hpp.read_text()
    out = {}
    for fn in ("sin", "cos"):
        start = text.index(f"struct remap_comp<N, ops::{fn}<T>, T, T>")
        body = text[start:text.index("};\n};", start)]
        flt, _, dbl = body.partition("else")
        for kind, part in (("f32", flt), ("f64", dbl)):
            ks = re.findall(r"DECLK\(k(\d+),\s*([-+0-9.eE]+)(f?)\)", part)
            assert ks and [int(i) for i, _, _ in ks] == list(range(1, len(ks) + 1)), (fn, kind)
            assert all((suf == "f") == (kind == "f32") for _, _, suf in ks), (fn, kind)
            out[(fn, kind)] = [R.round_to(Fraction(v), kind) for _, v, _ in ks]
    return out


def _fma(a, b, c, kind):
    return R.round_to(a * b + c, kind)


def _nest(x2, ks, kind):
    if len(ks) == 2:
        return _fma(x2, ks[1], ks[0], kind)
    return _fma(_nest(x2, ks[1:], kind), x2, ks[0], kind)


def poly(fn, x: Fraction, ks, kind):
    x2 = R.round_to(x * x, kind)
    p = _nest(x2, ks, kind)
    return R.round_to(x * p, kind) if fn == "sin" else p


def inputs(kind, n):
    """Seeded points of [-pi, pi] (the polynomials' domain) plus edges."""
    rng = np.random.default_rng(SEED + (kind == "f64"))
    dt = np.float32 if kind == "f32" else np.float64
    pi = dt(np.pi)
    xs = rng.uniform(-float(pi), float(pi), n - 9).astype(dt)
    edges = np.array([0.0, -0.0, pi, -pi, pi / 2, -pi / 2, 1e-7, -1e-30, 1.0], dt)
    return np.concatenate([edges, xs]).astype(dt)


def to_bits(v, kind):
    return R.bits(v, kind)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", type=Path, default=Path("/root/reference"))
    a = ap.parse_args(argv)
    tables = coefficient_tables(a.ref / "src" / "vm" / "simd.hpp")
    doc = {"how": "reference sin/cos polynomials (simd.hpp:955-1100) with coefficients read "
                  "from its source, nested fma order of detail::nest_fma (simd.hpp:193-201), "
                  "exact rational arithmetic rounded once per operation "
                  "(tests/golden/make_sincos.py)"}
    for kind, n in (("f32", N_F32), ("f64", N_F64)):
        xs = inputs(kind, n)
        xb = xs.view(np.uint32 if kind == "f32" else np.uint64)
        row = {"x_bits": [int(v) for v in xb]}
        for fn in ("sin", "cos"):
            ys = []
            for x in xs:
                v = poly(fn, Fraction(float(x)), tables[(fn, kind)], kind)
                if v == 0:
                    # an exact zero keeps IEEE's sign: sin(+-0) = +-0 * k1 with k1 > 0
                    v = np.copysign(0.0, float(x)) if fn == "sin" else 0.0
                    ys.append(int(np.array(v, dtype=xs.dtype).view(xb.dtype)))
                else:
                    ys.append(to_bits(v, kind))
            row[f"{fn}_bits"] = ys
        row["sha256"] = hashlib.sha256(json.dumps(row, sort_keys=True).encode()).hexdigest()
        doc[kind] = row
    (HERE / "sincos.json").write_text(json.dumps(doc) + "\n")
    print("sincos", {k: len(doc[k]["x_bits"]) for k in ("f32", "f64")}, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
