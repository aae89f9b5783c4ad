"""The oracle pinned against closed forms, brute force and exact arithmetic.

The reference commits no golden vectors; its tests pin kernels with
closed-form validators (reference tests/programs.hpp:288-389) and sorter
invariants.  These tests pin the CPU oracle the same way before it is used
as the checker of the GPU backend.
"""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

from loom_amd import workloads
from tests import oracle_lib as O

M64 = (1 << 64) - 1


def fib(n):
    a, b = 0, 1
    for _ in range(n):
        a, b = b, a + b
    return a


def collatz(x):
    c = 0
    while x > 1:
        x = x // 2 if x % 2 == 0 else 3 * x + 1
        assert x < 1 << 64
        c += 1
    return c


def test_fib_all_inputs():
    n = np.arange(91, dtype=np.uint64)
    (c,), err = O.run("fib_u64", [n])
    assert err == 0
    assert [int(v) for v in c] == [fib(i) for i in range(91)]


def test_fib_u32_wraps():
    n = np.arange(60, dtype=np.uint32)
    (c,), _ = O.run("fib_u32", [n])
    assert [int(v) for v in c] == [fib(i) % (1 << 32) for i in range(60)]


def test_collatz_brute_force():
    x = workloads.collatz_inputs(3000)
    x = np.concatenate([x, np.array([1, 2, 3, 27, 97, 871, (1 << 32) - 1], np.uint64)])
    (s,), err = O.run("collatz_u64", [x])
    assert err == 0
    assert [int(v) for v in s] == [collatz(int(v)) for v in x]
    assert int(s[-4]) == 111 and int(s[-3]) == 118 and int(s[-2]) == 178


def test_factorial():
    a = np.arange(30, dtype=np.uint64)
    (r,), _ = O.run("factorial_u64", [a])
    assert [int(v) for v in r] == [(math.factorial(i) & M64) if i > 1 else 1 for i in range(30)]


def test_jump4_and_out_of_range_index():
    i = np.array([0, 1, 2, 3, 4, 5, 1 << 31], np.uint32)
    (o,), _ = O.run("jump4_u32", [i])
    assert list(map(int, o)) == [1, 3, 7, 11, 4, 5, 1 << 31]


def int_mix_expected(a, b, bits):
    """int_mix with the op contract: modular arithmetic, x/0 = all ones,
    x%0 = x, shift counts modulo the promoted width (32 or 64)."""
    m = (1 << bits) - 1
    w = 32 if bits <= 32 else 64
    s = (a + b) & m
    d = (a - b) & m
    mul = (a * b) & m
    q = m if b == 0 else a // b
    r = a if b == 0 else a % b
    x1, x2, x3, x4 = a & b, a | b, a ^ b, (~a) & m
    l = ((a << (b % w)) & ((1 << w) - 1)) & m
    h = a >> (b % w)
    mn, mx = min(a, b), max(a, b)
    ng = (-a) & m
    la, lo, ln = int(bool(a) and bool(b)), int(bool(a) or bool(b)), int(not a)
    c1, c2, c3 = int(a < b), int(a >= b), int(a != b)
    sc = [10, 0x20, 3, 4][c1]
    p1 = int(bool(c2) and a < b)
    p2 = int(bool(c1) or a >= b)
    x1 = x1 ^ x2 ^ x3 ^ x4
    s = (s + l + h) & m
    d = (d + mn + mx) & m
    mul = (mul + ng) & m
    q = (q + r) & m
    la = (la + lo + ln + c1 + c2 + c3 + sc + p1 + p2) & m
    return [s, d, mul, q, x1, la]


@pytest.mark.parametrize("name,bits", [("int_mix_u64", 64), ("int_mix_u32", 32)])
def test_int_mix(name, bits):
    dt = np.uint64 if bits == 64 else np.uint32
    raw = workloads.splitmix64(7, 4000)
    a = (raw & np.uint64((1 << bits) - 1)).astype(dt)
    b = (workloads.splitmix64(8, 4000) & np.uint64((1 << bits) - 1)).astype(dt)
    # small and edge operands: zero divisors, shift counts around the width
    b[:200] = np.arange(200) % 70
    a[:5] = [0, 1, (1 << bits) - 1, 1 << (bits - 1), 12345]
    outs, err = O.run(name, [a, b])
    assert err == 0
    for i in range(len(a)):
        exp = int_mix_expected(int(a[i]), int(b[i]), bits)
        got = [int(o[i]) for o in outs]
        assert got == exp, (i, int(a[i]), int(b[i]), got, exp)


# ---- exact float32 arithmetic for the fp_mix check ---------------------------------
def f32(x):
    return float(np.float32(x))


def round_f32(q: Fraction) -> float:
    """Correctly rounded (nearest-even) float32 of a rational."""
    if q == 0:
        return 0.0
    x = float(q)  # nearest double, then fix the float32 rounding exactly
    c = np.float32(x)
    lo, hi = np.nextafter(c, np.float32(-np.inf)), np.nextafter(c, np.float32(np.inf))
    best = min((lo, c, hi), key=lambda v: (abs(Fraction(float(v)) - q),
                                           struct.unpack("<I", struct.pack("<f", v))[0] & 1))
    return float(best)


def fma32(x, y, z):
    return round_f32(Fraction(x) * Fraction(y) + Fraction(z))


def test_fp_mix_f32_exact():
    rng = np.random.default_rng(5)
    x = rng.uniform(-900, 900, 800).astype(np.float32)
    y = rng.uniform(-50, 50, 800).astype(np.float32)
    y[::7] = np.float32(0.25)
    x[:4] = [0.0, -0.0, 1e-3, 4e9]
    outs, err = O.run("fp_mix_f32", [x, y])
    assert err == 0
    a_o, b_o, cs_o, sn_o, u_o, sum_o = outs
    for i in range(len(x)):
        X, Y = float(x[i]), float(y[i])
        a = f32(f32(X * Y) + X)
        b = f32(X / Y)
        c = fma32(X, Y, b)
        d = fma32(X, Y, -a)
        g1 = f32(math.sqrt(abs(X)))
        g2 = f32(1.0 / f32(math.sqrt(g1))) if g1 != 0 else math.inf
        r1 = f32(math.copysign(math.floor(abs(b) + 0.5), b) + math.floor(b))
        r3 = f32(math.ceil(b) + math.trunc(b))
        mn = Y if Y < X else X
        mx = Y if X < Y else X
        cs = math.copysign(abs(X), Y)
        sel = X < Y
        pick = mn if sel else mx
        k = [0.5, 1.25, -300.0][int(sel)]
        a = f32(f32(a + c) + d)
        b = f32(f32(f32(b + g2) + r1) + r3)
        cs = f32(f32(cs + pick) + k)
        u = 0 if X <= 0 else min(int(X), (1 << 32) - 1)
        s = f32(a + b)
        assert float(a_o[i]) == a, i
        assert float(b_o[i]) == b or (math.isnan(b) and math.isnan(b_o[i])), i
        assert float(cs_o[i]) == cs, i
        assert int(u_o[i]) == u, i
        assert float(sum_o[i]) == s or (math.isnan(s) and math.isnan(sum_o[i])), i
        ang = f32(X * f32(0.001))
        assert abs(float(sn_o[i]) - math.sin(ang) * math.cos(ang)) < 2e-6 or abs(ang) > math.pi


# ---- sin/cos polynomials ---------------------------------------------------------
def test_sin_cos_polynomials_close_to_libm():
    L = O.lib()
    xs = np.linspace(-math.pi, math.pi, 20001)
    e32 = max(max(abs(L.oracle_sin_f32(float(np.float32(x))) - math.sin(float(np.float32(x)))),
                  abs(L.oracle_cos_f32(float(np.float32(x))) - math.cos(float(np.float32(x)))))
              for x in xs)
    e64 = max(max(abs(L.oracle_sin_f64(x) - math.sin(x)), abs(L.oracle_cos_f64(x) - math.cos(x)))
              for x in xs)
    assert e32 < 1e-6  # float32 evaluation; measured 4.8e-7
    assert e64 < 2e-15  # measured 7.4e-16
    assert L.oracle_sin_f32(0.0) == 0.0 and L.oracle_cos_f64(0.0) == 1.0


# ---- literals --------------------------------------------------------------------------
@pytest.mark.parametrize("text,value", [
    ("0", 0.0), ("1", 1.0), ("0.5", 0.5), ("-0.5", -0.5), ("0.25", 0.25), ("1.25", 1.25),
    ("-3e2", -300.0), ("1e3", 1000.0), ("100", 100.0), ("+2", 2.0), ("16.5", 16.5),
    ("1e", 1.0),      # empty exponent is ignored (string.hpp:343)
    ("e1", 0.0),      # no digits before the exponent: a literal 0 (string.hpp:330)
    (".", 0.0),
])
def test_literal_exact_cases(text, value):
    assert O.parse("f32", text) == value
    assert O.parse("f64", text) == value


@pytest.mark.parametrize("text", ["x", "-", "", "1.x", "cmp", "0x1"])
def test_literal_rejects_names(text):
    assert O.parse("f32", text) is None


def test_literal_inexact_within_ulps_of_strtof():
    # the reference parser is not correctly rounded; it stays within a few
    # ulp of the correctly rounded value for the literals kernels use
    for t in ["0.1", "0.093073", "6.283185307179586", "0.0001", "1e-12", "1e-6",
              "3.14159265", "0.999", "81.6", "1e5", "681.33", "0.042612", "0.5135",
              "295.6", "1.5", "0.75", "12"]:
        got = np.float32(O.parse("f32", t))
        want = np.float32(float(t))
        ulps = abs(int(got.view(np.int32)) - int(want.view(np.int32)))
        assert ulps <= 4, (t, got, want)
        got64 = O.parse("f64", t)
        ulps64 = abs(int(np.float64(got64).view(np.int64)) - int(np.float64(float(t)).view(np.int64)))
        assert ulps64 <= 8, (t, got64)


# the reference's own checks of its float parser (src/tests/string.cpp:40-56):
# the parsed value equals the compiler's (correctly rounded) literal
REFERENCE_S2F_CHECKS = [("f32", t) for t in ("0", "1", "1234.567", "1e2", "0.5e2", "0.5e-2",
                                             "1e-8")] + \
                       [("f64", t) for t in ("0", "-1234.567", "-1e2", "-0.5e2", "-0.5e-2",
                                             "-2430.5e4")]


@pytest.mark.parametrize("kind,text", REFERENCE_S2F_CHECKS)
def test_literal_reference_checks(kind, text):
    from tests import literal_ref as R
    got = O.parse(kind, text)
    assert got is not None
    assert float(got) == float(R.correctly_rounded(text, kind)), (kind, text, got)
    # and the independent restatement agrees with the shared C++ one
    assert float(got) == float(R.s2f(text, kind))


@pytest.mark.parametrize("text", ["e334g", "-", "+"])
def test_literal_reference_rejects(text):
    from tests import literal_ref as R
    assert O.parse("f32", text) is None
    assert R.s2f(text, "f32") is None


def _int_literal(text, dtype):
    """Integer literal value as the reference's s2i_base_prefix_fn reads it
    (string.hpp:286-306): optional sign for signed types, 0x / 0b prefixes."""
    signed = np.issubdtype(dtype, np.signedinteger)
    neg = False
    if signed and text[:1] in "+-":
        neg, text = text[0] == "-", text[1:]
    base = 10
    if text[:2] in ("0x", "0X"):
        base, text = 16, text[2:]
    elif text[:2] in ("0b", "0B"):
        base, text = 2, text[2:]
    v = int(text, base)
    bits = 8 * np.dtype(dtype).itemsize
    return (-v if neg else v) % (1 << bits)


def _check_kernel_literals(path):
    from tests import literal_ref as R
    seen = 0
    for name in O.kernel_names(path):
        for text, dt, bits in O.literals(name, path):
            if dt in (np.float32, np.float64):
                kind = "f32" if dt == np.float32 else "f64"
                want = R.bits(R.s2f(text, kind), kind)
            else:
                want = _int_literal(text, dt)
            assert bits == want, (name, text, dt, hex(bits), hex(want))
            seen += 1
    return seen


def test_every_kernel_literal_matches_restatement():
    """Every literal operand of every kernel of the library (the constants
    the GPU computes with come from the same lowering) against the Python
    restatement of the reference's parser."""
    assert _check_kernel_literals(O.PATH) > 300


def test_reference_kernel_literals_match_restatement():
    """The same over the reference's own kernels (path tracers, test
    programs) where they were built from its checkout."""
    if not O.REF_PATH.exists():
        pytest.skip("reference kernels not built (no checkout)")
    assert _check_kernel_literals(O.REF_PATH) > 500


def test_integer_literals():
    assert O.parse("u64", "0x20") == 32
    assert O.parse("u64", "0b11") == 3
    assert O.parse("u64", "18446744073709551615") == M64
    assert O.parse("u64", "-1") is None          # unsigned: no sign accepted
    assert O.parse("i32", "-0x10") == -16
    assert O.parse("u64", "0x") is None


# ---- path tracer: IR (interpreted) vs direct C++ -------------------------------------
def test_smallpt9_ir_matches_direct_cpp():
    """The IR kernel run by the interpreter equals the path tracer written
    directly in C++ bit for bit, on a whole 48x32x8 frame and on samples of
    the 1024x768x64 frame (camera corners, centre)."""
    w, h, spp = 48, 32, 8
    ins = workloads.pt_inputs(w, h, spp)
    ir, err = O.run("smallpt9_f32", ins)
    assert err == 0
    d = O.smallpt9_direct(*ins)
    for a, b in zip(ir, d):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    job = workloads.pt_job(1024, 768, 64)
    sid = np.concatenate([np.arange(0, 4096), np.arange(25_000_000, 25_004_096),
                          np.arange(1024 * 768 * 64 - 4096, 1024 * 768 * 64)]).astype(np.uint32)
    jobs = np.full(len(sid), job, np.uint32)
    ir, _ = O.run("smallpt9_f32", [sid, jobs])
    d = O.smallpt9_direct(sid, jobs)
    for a, b in zip(ir, d):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_smallpt9_image_invariants():
    """Radiance is finite and non-negative; the emitter is visible; the
    image is not black and bounded by the emission (12 per channel)."""
    w, h, spp = 40, 30, 16
    img = workloads.pt_image(*O.smallpt9_direct(*workloads.pt_inputs(w, h, spp)), w, h, spp)
    assert np.isfinite(img).all() and (img >= 0).all()
    assert img.max() <= 12.0 and 0.05 < img.mean() < 5.0
    # the ceiling light (top rows) is brighter than the floor (bottom rows)
    assert img[-3:].mean() > img[:3].mean()


# ---- Threefry-2x32 and cubic bisection -------------------------------------------------
def threefry_ref(c0, c1, k0=0xDEAD, k1=0xBEEF):
    """Threefry-2x32, 32 rounds, from its published definition (Salmon et al.,
    SC 2011): rotations 13 15 26 6 17 29 16 24, key words k0, k1 and
    0x1BD11BDA ^ k0 ^ k1, an injection after every four rounds."""
    M = np.uint64(0xFFFFFFFF)
    ks = [np.uint64(k0), np.uint64(k1), np.uint64(0x1BD11BDA ^ k0 ^ k1)]
    rot = [13, 15, 26, 6, 17, 29, 16, 24]
    x0 = (c0.astype(np.uint64) + ks[0]) & M
    x1 = (c1.astype(np.uint64) + ks[1]) & M
    for r in range(32):
        x0 = (x0 + x1) & M
        s = np.uint64(rot[r % 8])
        x1 = ((x1 << s) | (x1 >> (np.uint64(32) - s))) & M
        x1 = x1 ^ x0
        if r % 4 == 3:
            j = r // 4 + 1
            x0 = (x0 + ks[j % 3]) & M
            x1 = (x1 + ks[(j + 1) % 3] + np.uint64(j)) & M
    return x0.astype(np.uint32), x1.astype(np.uint32)


def test_threefry_matches_definition():
    c0, c1 = workloads.threefry_inputs(5000)
    c0[:4] = [0, 0xFFFFFFFF, 0xFFFFFFFF - 39916800, 123]
    c1[:4] = [0, 0xFFFFFFFF, 7, 0]
    (x0, x1, n0, n1), err = O.run("threefry2x32_u32", [c0, c1])
    assert err == 0
    w0, w1 = threefry_ref(c0, c1)
    assert np.array_equal(x0, w0) and np.array_equal(x1, w1)
    ctr = (c1.astype(np.uint64) << np.uint64(32)) | c0.astype(np.uint64)
    nxt = ctr + np.uint64(39916801)  # wraps modulo 2^64
    assert np.array_equal(n0, (nxt & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert np.array_equal(n1, (nxt >> np.uint64(32)).astype(np.uint32))


def _round32(q: Fraction) -> np.float32:
    """q rounded to float32, nearest with ties to even (exact)."""
    f = np.float32(float(q))
    best = f
    for cand in (np.nextafter(f, np.float32(-np.inf)), np.nextafter(f, np.float32(np.inf))):
        dc, db = abs(Fraction(float(cand)) - q), abs(Fraction(float(best)) - q)
        if dc < db or (dc == db and int(cand.view(np.uint32)) % 2 == 0):
            best = cand
    return np.float32(best)


def _fma(dt, x, y, z):
    q = Fraction(float(x)) * Fraction(float(y)) + Fraction(float(z))
    return _round32(q) if dt == np.float32 else np.float64(float(q))


def bisect_ref(dt, a, b, c, d, err_lit):
    """The bisection of the reference's validator (tests/root_finding.hpp:
    193-216) in exact IEEE arithmetic of type dt, with the loop caps of the
    kernel (40 widenings, 256 halvings)."""
    a, b, c, d = (dt(v) for v in (a, b, c, d))

    def cubic(x):
        x2 = dt(x * x)
        x3 = dt(x2 * x)
        return _fma(dt, a, x3, _fma(dt, b, x2, _fma(dt, c, x, d)))

    def split(p, q):
        return (cubic(p) < 0) != (cubic(q) < 0)

    lo, hi, w = dt(-10), dt(10), 0
    while True:
        s = split(lo, hi)
        w += 1
        if s or w >= 40:
            break
        lo, hi = dt(lo * dt(10)), dt(hi * dt(10))
    mid = dt(dt(dt(hi - lo) * dt(0.5)) + lo)
    fm, i = abs(cubic(mid)), 0
    if fm >= err_lit:
        while True:
            left = split(lo, mid)
            hi, lo = (mid, lo) if left else (hi, mid)
            mid = dt(dt(dt(hi - lo) * dt(0.5)) + lo)
            fm = abs(cubic(mid))
            i += 1
            if not (fm >= err_lit and i < 256):
                break
    return mid, i


@pytest.mark.parametrize("name,dt", [("cubic_bisect_f32", np.float32),
                                     ("cubic_bisect_f64", np.float64)])
def test_cubic_bisect_exact(name, dt):
    ins = workloads.cubic_inputs(120, dt)
    (mid, it), err = O.run(name, ins)
    assert err == 0
    err_lit = dt(O.parse("f32" if dt == np.float32 else "f64", "0.1"))
    for k in range(len(ins[0])):
        m, i = bisect_ref(dt, *(col[k] for col in ins), err_lit)
        assert mid[k].tobytes() == dt(m).tobytes() and int(it[k]) == i, (k, mid[k], m, it[k], i)
        assert i < 256  # every input of the workload converges before the cap
