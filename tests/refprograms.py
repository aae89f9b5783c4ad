"""The reference's own test programs (src/tests/programs.hpp), compiled from
the reference checkout into oracle/_ref/ (oracle/build_ref.py), and the
inputs its test suites draw for them: shared by test_refprograms.py (CPU)
and test_gpu_refprograms.py."""
from __future__ import annotations

import ctypes as C

import numpy as np

from tests import oracle_lib as O

# programs.hpp kernel -> (input generator, checked by: "validator" = the
# reference's fn::val_* / std functor, "invariant" = every output non-zero,
# "fuzzy" = its reference function within 100 ULP on 99% of inputs)
# Generators follow src/tests/generators.hpp:30-210 and the suites that pick
# them (src/tests/vm.cpp:262-325, vm_simd_compact.cpp:180-251).
PROGRAMS = {
    "increment": ("incr", "validator"),      # GM_Incr: [lowest, max - 1]
    "fac_test": ("fac", "validator"),        # GM_FAC: [0, 12] / [0, 20]
    "fib_test": ("fib", "validator"),        # GM_FIB: [0, 80]
    "union_test": ("canon", "validator"),    # GMCanon: [0, 1)
    "union_cmov_test": ("canon", "validator"),
    "jmp4_test": ("u3", "validator"),        # GM_U3: [0, 3]
    "lshift_test": ("shift", "validator"),   # GM_Shift: two of [2, 10]
    "rshift_test": ("shift", "validator"),
    "bit_and_test": ("bits2", "validator"),  # GM_Bit_Bin: two of [0, max]
    "bit_or_test": ("bits2", "validator"),
    "bit_xor_test": ("bits2", "validator"),
    "bit_not_test": ("bits1", "validator"),  # GM_0U
    "bool_and_test": ("bool2", "validator"),  # GM_Bool_Bin
    "bool_or_test": ("bool2", "validator"),
    "bool_not_test": ("bool1", "validator"),  # GM_Bool
    "sct_inv_test": ("u3", "invariant"),
    "slt_inv_test": ("u3", "invariant"),
    "abs_test": ("float", "invariant"),       # GMFloat: [-1e6, 1e6]
    "trunc_test": ("float", "invariant"),
    "ceilfloor": ("float", "invariant"),
    "minmax_test": ("float", "invariant"),    # GM_F2
    "arithmetic_test": ("float", "invariant"),  # GM_F3
    "sin_cos_test": ("trig", "invariant"),    # GMTrig: [-pi, pi]
    "sqrt_test": ("positive", "invariant"),   # GMFloatPositive: [1e-6, 1e6]
    "rsqrt_test": ("positive", "invariant"),
    # src/benchmarks/fp_mix.hpp against its fp_mix_ref within 100 ULP on 99%
    # of GM_F3 inputs (src/tests/vm_fp_mix.cpp:33-41)
    "fp_mix_k": ("float", "fuzzy"),
}
TYPES = {"i8": np.int8, "i16": np.int16, "i32": np.int32, "i64": np.int64, "u8": np.uint8,
         "u16": np.uint16, "u32": np.uint32, "u64": np.uint64, "f32": np.float32,
         "f64": np.float64}
# every instantiation oracle/ref_kernels.hpp registers
NAMES = ([f"ref_increment_{t}" for t in ("i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64")]
         + ["ref_fac_test_u32", "ref_fac_test_u64", "ref_fib_test_u64"]
         + [f"ref_{p}_{t}" for p in ("union_test", "union_cmov_test") for t in ("f32", "f64")]
         + [f"ref_jmp4_test_{t}" for t in ("u8", "u16", "u32", "u64", "f32")]
         + [f"ref_{p}_{t}" for p in ("lshift_test", "rshift_test") for t in ("u16", "u32", "u64")]
         + [f"ref_{p}_{t}" for p in ("bit_and_test", "bit_or_test", "bit_xor_test",
                                     "bit_not_test", "bool_and_test", "bool_or_test",
                                     "bool_not_test", "sct_inv_test", "slt_inv_test")
            for t in ("u8", "u16", "u32", "u64")]
         + [f"ref_{p}_{t}" for p in ("abs_test", "trunc_test", "ceilfloor", "minmax_test",
                                     "arithmetic_test", "sin_cos_test", "sqrt_test",
                                     "rsqrt_test", "fp_mix_k") for t in ("f32", "f64")])


def split(name):
    """("fib_test", "u64") from "ref_fib_test_u64"."""
    body, ty = name[len("ref_"):].rsplit("_", 1)
    return body, ty


def inputs(name, n, seed=0):
    """Input columns of program `name` for `n` threads, drawn from the
    reference's generator ranges for it."""
    body, ty = split(name)
    gen, _ = PROGRAMS[body]
    _, itys, _ = O.io(name, O.REF_PATH)
    rng = np.random.default_rng([seed, sum(map(ord, name))])
    cols = []
    for dt in itys:
        dt = np.dtype(dt)
        info = np.iinfo(dt) if dt.kind in "iu" else None
        if gen == "incr":
            c = rng.integers(info.min, info.max, n, dtype=dt, endpoint=False)
        elif gen == "fac":
            c = rng.integers(0, 20 if dt.itemsize == 8 else 12, n, dtype=dt, endpoint=True)
        elif gen == "fib":
            c = rng.integers(0, 80, n, dtype=dt, endpoint=True)
        elif gen == "canon":
            c = rng.random(n).astype(dt)
        elif gen == "u3":
            c = rng.integers(0, 3, n, dtype=dt, endpoint=True)
        elif gen == "shift":
            c = rng.integers(2, 10, n, dtype=dt, endpoint=True)
        elif gen in ("bits1", "bits2"):
            c = rng.integers(0, info.max, n, dtype=dt, endpoint=True)
        elif gen in ("bool1", "bool2"):
            c = rng.integers(0, 1, n, dtype=dt, endpoint=True)
        elif gen == "float":
            c = rng.uniform(-1e6, 1e6, n).astype(dt)
        elif gen == "trig":
            c = rng.uniform(-np.pi, np.pi, n).astype(dt)
        elif gen == "positive":
            c = rng.uniform(1e-6, 1e6, n).astype(dt)
        else:
            raise KeyError(gen)
        cols.append(c)
    # every range edge the generator allows, first
    edges = {"fac": [0, 1, 2, 12], "fib": [0, 1, 2, 80], "u3": [0, 1, 2, 3], "shift": [2, 10]}
    if gen in edges and n >= len(edges[gen]):
        for c in cols:
            c[:len(edges[gen])] = edges[gen]
    return cols


def validate(name, cols):
    """The reference validator's outputs for `cols`, or None for a program
    checked by an invariant."""
    body, _ = split(name)
    if PROGRAMS[body][1] == "invariant":
        return None
    L = O.lib(O.REF_PATH)
    L.ref_validate.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.c_void_p, C.c_uint64]
    _, _, otys = O.io(name, O.REF_PATH)
    assert len(otys) == 1
    cols = [np.ascontiguousarray(c) for c in cols]
    out = np.zeros(len(cols[0]), otys[0])
    ip = (C.c_void_p * 16)(*[c.ctypes.data for c in cols])
    assert L.ref_validate(name.encode(), ip, out.ctypes.data, len(out)) == 0, name
    return out


def ulp_mismatches(a, b, ulp=100):
    """How many pairs the reference's loom::almost_equal(a, b, ulp) rejects."""
    L = O.lib(O.REF_PATH)
    L.ref_ulp_mismatches.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint, C.c_int]
    L.ref_ulp_mismatches.restype = C.c_uint64
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.dtype == b.dtype and a.dtype.kind == "f" and len(a) == len(b)
    return L.ref_ulp_mismatches(a.ctypes.data, b.ctypes.data, len(a), ulp, a.dtype.itemsize)


# the reference's random-number IR macros (oracle/ref_kernels.hpp rng_macros)
RNG = "ref_rng_macros_f32"
# counter words that overflow when bumped by 39916801, and their neighbours
RNG_EDGES = [(0, 0), (0xFFFFFFFF, 0), (0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF - 39916800, 7),
             (0xFFFFFFFF - 39916801, 7), (12345, 0xFFFFFFFF)]


def rng_inputs(n, seed=0):
    rng = np.random.default_rng([seed, 7])
    c0 = rng.integers(0, 2**32, n, dtype=np.uint32)
    c1 = rng.integers(0, 2**32, n, dtype=np.uint32)
    for i, (a, b) in enumerate(RNG_EDGES[:n]):
        c0[i], c1[i] = a, b
    return [c0, c1]


def rng_reference(c0, c1):
    """(w0, w1, n0, n1, f0, f1) from the reference's C++ functions."""
    L = O.lib(O.REF_PATH)
    L.ref_rng.argtypes = [C.c_void_p] * 8 + [C.c_uint64]
    L.ref_rng.restype = None
    c0, c1 = np.ascontiguousarray(c0, np.uint32), np.ascontiguousarray(c1, np.uint32)
    n = len(c0)
    outs = [np.zeros(n, np.uint32) for _ in range(4)] + [np.zeros(n, np.float32) for _ in range(2)]
    L.ref_rng(c0.ctypes.data, c1.ctypes.data, *[o.ctypes.data for o in outs], n)
    return outs


JMP4 = [n for n in NAMES if n.startswith("ref_jmp4_test_")]


def jmp4_wide_inputs(name, n=4096, seed=0):
    """Jump indexes over the whole range of the index type (0-255 first):
    indexes past the four labels end the thread (the op contract), leaving
    the index as the output, which is what the reference's validator returns
    for them (src/tests/programs.hpp:325-337)."""
    _, itys, _ = O.io(name, O.REF_PATH)
    dt = np.dtype(itys[0])
    rng = np.random.default_rng([seed, 4])
    c = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    c[:256] = np.arange(256, dtype=dt)
    return [c]
