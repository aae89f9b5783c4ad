"""The reference's path-tracing and vector macro checks (test
infrastructure): input generators and the comparison of
src/tests/vm_smallpt.cpp:311-700 and src/tests/vm_linal.cpp:68-140, for the
kernels of oracle/ref_macros.hpp.

Inputs follow the reference's generator ranges (vm_smallpt.cpp:323-335:
floats in [-1e6, 1e6], positive floats in [0.1, 1e6], [-1, 1], canonical
[0, 1), indices of refraction in [1, 10], film sizes in [30, 1000]; rays
from a random origin, half of them aimed at the primitive), drawn from this
repository's seeded generator.  A row passes when its outputs agree within
the reference's ULP bound (100 for the path tracer, 30 for vector algebra)
by its almost_equal, an intersection when both miss or both hit at a
distance within the bound; a check passes when at most 1% of rows fail
(2% for microfacet sampling), the reference's fuzzy-test rule."""
from __future__ import annotations

import ctypes as C

import numpy as np

from loom_amd import workloads
from tests import oracle_lib as O

F32 = np.float32


class Gen:
    def __init__(self, seed: int, n: int):
        self.seed = seed
        self.n = n

    def _u(self):
        self.seed += 1
        r = workloads.splitmix64(self.seed, self.n)
        return (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def f(self, lo=-1e6, hi=1e6):
        return (lo + (hi - lo) * self._u()).astype(F32)

    def canon(self):
        return np.minimum(self._u().astype(F32), F32(np.nextafter(F32(1), F32(0))))

    def bools(self):
        return self._u() < 0.5

    def ints(self, lo, hi):
        return (lo + np.floor(self._u() * (hi - lo + 1))).astype(np.uint32)

    def v3(self, lo=-1e6, hi=1e6):
        return np.stack([self.f(lo, hi) for _ in range(3)])

    def unit3(self):
        return normalize(self.v3())


def normalize(v):
    v = v.astype(F32)
    length = np.sqrt((v * v).sum(axis=0, dtype=F32)).astype(F32)
    return (v / length).astype(F32)


def cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]]).astype(F32)


def rays(g: Gen, target):
    """Origins uniform in the float range; half the directions toward
    `target` (a point on the primitive), half random."""
    o = g.v3()
    aim = g.bools()
    d = np.where(aim, normalize(target - o), g.unit3())
    return list(o) + list(d)


def _disk(g):
    p, n, r = g.v3(), g.unit3(), g.f()
    return rays(g, p) + list(p) + list(n) + [r]


def _sphere(g):
    c, r = g.v3(), g.f(0.1, 1e6)
    return rays(g, c) + list(c) + [r]


def _cylinder(g):
    p, r, h = g.v3(), g.f(0.1, 1e6), g.f(0.1, 1e6)
    side = np.stack([p[0] + r, p[1] + h * F32(0.5), p[2]]).astype(F32)
    return rays(g, side) + list(p) + [r, h]


def _triangle(g):
    v = [g.v3() for _ in range(3)]
    return rays(g, ((v[0] + v[1] + v[2]) / F32(3)).astype(F32)) + [c for x in v for c in x]


def _quad(g):
    v00, v01, v11 = g.v3(), g.v3(), g.v3()
    v10 = (v00 + ((v01 + v11) * F32(0.5) - v00) * F32(2)).astype(F32)
    mid = ((v00 + v11) * F32(0.5)).astype(F32)
    return rays(g, mid) + [c for x in (v00, v01, v11, v10) for c in x]


def _tangents(g, n):
    tu = cross(n, g.unit3())
    return tu, cross(n, tu)


def _rough_i(g):
    wi, n = g.unit3(), g.unit3()
    tu, tv = _tangents(g, n)
    return list(wi) + list(tu) + list(tv) + [g.canon(), g.canon()]


def _rough_n(g):
    wi, n = g.unit3(), g.unit3()
    n = np.where((n * wi).sum(axis=0) < 0, -n, n).astype(F32)
    wm = normalize(wi + n)
    tu, tv = _tangents(g, n)
    return list(wm) + list(n) + list(tu) + list(tv) + [g.canon(), g.canon()]


def _sample_space(g):
    w, h = g.ints(30, 1000), g.ints(30, 1000)
    x = np.floor(g.canon() * w).astype(np.uint32)
    y = np.floor(g.canon() * h).astype(np.uint32)
    return [x, y, w, h]


# name -> (input generator, kind, ulp, allowed failure fraction)
CHECKS = {
    "pm_disk": (_disk, "isect", 100, 0.01),
    "pm_quadratic": (lambda g: [g.f(), g.f(), g.f()], "isect", 100, 0.01),
    "pm_sphere": (_sphere, "isect", 100, 0.01),
    "pm_cylinder": (_cylinder, "isect", 100, 0.01),
    "pm_triangle": (_triangle, "isect", 100, 0.01),
    "pm_quad": (_quad, "isect", 100, 0.01),
    "pm_onb": (lambda g: list(g.unit3()), "floats", 100, 0.01),
    "pm_hemisphere": (lambda g: list(g.unit3()) + [g.canon(), g.canon()], "floats", 100, 0.01),
    "pm_sample_space": (_sample_space, "floats", 100, 0.01),
    "pm_ray_dir": (lambda g: [g.f(-1, 1), g.f(-1, 1)], "floats", 100, 0.01),
    "pm_light": (lambda g: [g.canon(), g.canon()], "floats", 100, 0.01),
    "pm_schlick": (lambda g: [g.f(-1, 1), g.f(1, 10)], "floats", 100, 0.01),
    "pm_reflect": (lambda g: list(g.unit3()) + list(g.unit3()), "floats", 100, 0.01),
    "pm_refract": (lambda g: list(g.unit3()) + list(g.unit3()), "floats", 100, 0.01),
    "pm_gtr": (lambda g: [g.canon(), g.canon()], "floats", 100, 0.01),
    "pm_gaf": (lambda g: [g.canon(), g.canon()], "floats", 100, 0.01),
    "pm_slope": (lambda g: [g.canon(), g.canon(), g.canon()], "floats", 100, 0.01),
    "pm_microfacet": (lambda g: list(g.unit3()) + [g.canon() for _ in range(4)], "floats", 100,
                      0.02),
    "pm_rough_i": (_rough_i, "floats", 100, 0.01),
    "pm_rough_n": (_rough_n, "floats", 100, 0.01),
    "pm_cross": (lambda g: list(g.v3()) + list(g.v3()), "floats", 30, 0.01),
    "pm_mat_vec": (lambda g: [g.f() for _ in range(12)], "floats", 30, 0.01),
    "pm_norm": (lambda g: list(g.v3()), "floats", 30, 0.01),
}


def inputs(name: str, n: int, seed: int = 0x5EED):
    gen = CHECKS[name][0]
    g = Gen(seed + sum(name.encode()), n)
    return [np.ascontiguousarray(c) for c in gen(g)]


# kernels whose C++ side has the slope sampling's sine and cosine, which the
# op contract evaluates with the reference's SIMD polynomial (valid on
# [-pi, pi]) where the reference's C++ calls std::sin / std::cos
CONTRACT_TRIG = {"pm_slope", "pm_microfacet"}


def reference(name: str, ins, contract_trig: bool = True):
    """Expected outputs from the reference's C++ (or its restated formula);
    with `contract_trig` the slope sampling uses the op contract's sine and
    cosine."""
    L = O.lib(O.REF_PATH)
    L.ref_pm_contract_trig.argtypes = [C.c_int]
    L.ref_pm_contract_trig(1 if (contract_trig and name in CONTRACT_TRIG) else 0)
    _, itys, otys = O.io(name, O.REF_PATH)
    n = len(ins[0])
    outs = [np.zeros(n, t) for t in otys]
    ip = (C.c_void_p * 32)(*[a.ctypes.data for a in ins])
    op = (C.c_void_p * 16)(*[a.ctypes.data for a in outs])
    fn = L.ref_pm
    fn.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_uint64]
    assert fn(name.encode(), ip, op, n) == 0, name
    return outs


def ulp_rows(a_cols, b_cols, ulp):
    """Rows whose float columns all agree within `ulp` (reference almost_equal)."""
    L = O.lib(O.REF_PATH)
    n = len(a_cols[0])
    a = [np.ascontiguousarray(c, dtype=F32) for c in a_cols]
    b = [np.ascontiguousarray(c, dtype=F32) for c in b_cols]
    ap = (C.c_void_p * len(a))(*[c.ctypes.data for c in a])
    bp = (C.c_void_p * len(b))(*[c.ctypes.data for c in b])
    ok = np.zeros(n, np.uint8)
    fn = L.ref_ulp_rows
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_uint, C.c_void_p]
    fn(ap, bp, len(a), n, ulp, ok.ctypes.data)
    return ok.astype(bool)


def failures(name: str, got, want) -> float:
    """Fraction of rows failing the reference's comparison."""
    kind, ulp = CHECKS[name][1], CHECKS[name][2]
    if kind == "isect":
        gh, wh = got[0] != 0, want[0] != 0
        close = ulp_rows([got[1]], [want[1]], ulp)
        ok = np.where(gh & wh, close, gh == wh)
    else:
        ok = ulp_rows(got, want, ulp)
    return float(1.0 - ok.mean())
