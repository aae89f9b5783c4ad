"""A second, independent restatement of the reference's float literal parser
(test infrastructure).

The reference gives an IR literal such as "0.093073" the value its own
parser computes (src/common/string.hpp:312-352, `s2f_fn`), not the correctly
rounded one: the whole part converted to T, plus the fraction digits as an
integer times 10^-(digit count), times 10^exponent, times the sign, every
step rounded to T, with powers of ten by squaring in T
(ext/include/static_math/detail/cmath.inl:48-56, 618-628).

The product and the C++ oracle share one C++ restatement
(include/loom_vm/vm/literal.hpp).  This module restates the same rules in
Python with explicitly rounded IEEE arithmetic (every product, sum and
quotient is computed exactly as a rational and rounded once to binary32 or
binary64, round-to-nearest-even), so that a mistake in the C++ restatement
(a dropped step, a wrong power, a wrong rounding point) shows up as a
difference between the two.  It also gives the correctly rounded value of a
text, the criterion of the reference's own parser checks
(src/tests/string.cpp:33-56).
"""
from __future__ import annotations

import struct
from fractions import Fraction

# (mantissa bits incl. the hidden one, minimum normal exponent, maximum exponent)
_FORMATS = {"f32": (24, -126, 127), "f64": (53, -1022, 1023)}


def round_to(x: Fraction, kind: str) -> Fraction:
    """x rounded to the nearest value of the format (ties to even); overflow
    gives an infinity, represented here as None."""
    p, emin, emax = _FORMATS[kind]
    if x == 0:
        return Fraction(0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    # exponent e with 2^e <= a < 2^(e+1)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, emin)  # subnormals share the minimum exponent
    quantum = Fraction(2) ** (e - p + 1)
    q, r = divmod(a, quantum)
    q = int(q)
    if r * 2 > quantum or (r * 2 == quantum and q % 2 == 1):
        q += 1
    v = q * quantum
    if v >= Fraction(2) ** (emax + 1):
        return None
    return sign * v


def bits(v: Fraction | None, kind: str) -> int:
    """IEEE bit pattern of a representable value (None = +inf)."""
    f = float("inf") if v is None else float(v)  # exact: v is representable in binary64
    if kind == "f32":
        return struct.unpack("<I", struct.pack("<f", f))[0]
    return struct.unpack("<Q", struct.pack("<d", f))[0]


def _digits(s: str) -> int | None:
    if not s or any(c not in "0123456789" for c in s):
        return None
    return int(s)


def _mul(a, b, kind):
    return round_to(a * b, kind)


def _pow_sq(x: Fraction, e: int, kind: str) -> Fraction:
    """static_math pow_helper: x^e by squaring, each product rounded."""
    if e == 0:
        return Fraction(1)
    if e % 2 == 0:
        return _pow_sq(_mul(x, x, kind), e // 2, kind)
    return _mul(x, _pow_sq(_mul(x, x, kind), (e - 1) // 2, kind), kind)


def _pow10(e: int, kind: str) -> Fraction:
    if e == 0:
        return Fraction(1)
    if e > 0:
        return _pow_sq(Fraction(10), e, kind)
    return round_to(Fraction(1) / _pow_sq(Fraction(10), -e, kind), kind)


def s2f(text: str, kind: str) -> Fraction | None:
    """The reference parser's value of `text` in binary32 ("f32") or
    binary64 ("f64"); None when the text is not a number.  Follows the
    reference's decomposition: the first '.' and the first 'e'/'E' of the
    text split it into whole, fraction and exponent parts, empty parts are
    skipped, a sign is allowed in front of the text and of the exponent."""
    sign = 1
    if text[:1] in ("-", "+"):
        sign = -1 if text[0] == "-" else 1
        text = text[1:]
    if not text:
        return None
    sep = text.find(".")
    cands = [i for i in (text.find("e"), text.find("E")) if i >= 0]
    exp_at = min(cands) if cands else -1
    if sep >= 0:
        whole = text[:sep]
        # a marker before the '.' stays in `whole`, which then fails to parse
        frac = text[sep + 1:exp_at] if exp_at > sep else text[sep + 1:]
    else:
        whole = text[:exp_at] if exp_at >= 0 else text
        frac = ""
    expo = text[exp_at + 1:] if exp_at >= 0 else ""
    value = Fraction(0)
    if whole:
        d = _digits(whole)
        if d is None:
            return None
        value = round_to(Fraction(d), kind)
    if frac:
        d = _digits(frac)
        if d is None:
            return None
        term = _mul(round_to(Fraction(d), kind), _pow10(-len(frac), kind), kind)
        value = round_to(value + term, kind)
    if expo:
        es = 1
        if expo[0] in "-+":
            es = -1 if expo[0] == "-" else 1
            expo = expo[1:]
        d = _digits(expo)
        if d is None:
            return None
        value = _mul(value, _pow10(es * d, kind), kind)
    return value * sign


def correctly_rounded(text: str, kind: str) -> Fraction:
    """The value a C++ compiler gives the literal `text` (nearest, ties to
    even) — the criterion of the reference's parser checks."""
    return round_to(Fraction(text), kind)
