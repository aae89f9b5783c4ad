"""ctypes access to the CPU oracle (oracle/_build/liboracle.so): the checker
of the parity tests.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PATH = ROOT / "oracle" / "_build" / "liboracle.so"
# the reference's own kernels over the same interpreter (oracle/build_ref.py)
REF_PATH = ROOT / "oracle" / "_ref" / "liboracle_ref.so"
DTYPES = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.uint64, 4: np.int8, 5: np.int16,
          6: np.int32, 7: np.int64, 8: np.float32, 9: np.float64}
_libs = {}


def lib(path=PATH):
    path = Path(path)
    if path not in _libs:
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `python -m loom_amd.build`")
        L = C.CDLL(str(path))
        L.oracle_kernel_find.argtypes = [C.c_char_p]
        L.oracle_kernel_name.restype = C.c_char_p
        L.oracle_kernel_io.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int)]
        L.oracle_run.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                 C.c_uint64, C.c_uint64, C.c_int, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint32)]
        for nm, t in (("f32", C.c_float), ("f64", C.c_double)):
            getattr(L, f"oracle_parse_{nm}").argtypes = [C.c_char_p, C.POINTER(t)]
            for f in ("sin", "cos"):
                fn = getattr(L, f"oracle_{f}_{nm}")
                fn.argtypes = [t]
                fn.restype = t
        L.oracle_smallpt9_direct.argtypes = [C.c_void_p] * 5 + [C.c_uint64, C.c_uint64, C.c_int]
        L.oracle_parse_u64.argtypes = [C.c_char_p, C.POINTER(C.c_uint64)]
        L.oracle_collatz_direct.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
        L.oracle_kernel_literal.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_size_t,
                                            C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
        L.oracle_parse_i32.argtypes = [C.c_char_p, C.POINTER(C.c_int32)]
        L.oracle_trace.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_uint32,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32)]
        L.oracle_block_labels.argtypes = [C.c_int, C.c_int32, C.c_char_p, C.c_size_t]
        L.oracle_register_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.oracle_work.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.c_uint64, C.c_uint64, C.c_int,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]
        _libs[path] = L
    return _libs[path]


def io(name, path=PATH):
    L = lib(path)
    k = L.oracle_kernel_find(name.encode())
    if k < 0:
        raise KeyError(name)
    ni, no, nr = C.c_int(), C.c_int(), C.c_int()
    it, ot = (C.c_int * 32)(), (C.c_int * 32)()
    L.oracle_kernel_io(k, C.byref(ni), C.byref(no), it, ot, C.byref(nr))
    return k, [DTYPES[it[i]] for i in range(ni.value)], [DTYPES[ot[i]] for i in range(no.value)]


def run(name, inputs, threads=8, path=PATH):
    """Outputs of kernel `name` on host columns `inputs` (oracle)."""
    k, itys, otys = io(name, path)
    ins = [np.ascontiguousarray(a, dtype=t) for a, t in zip(inputs, itys)]
    n = len(ins[0])
    outs = [np.zeros(n, t) for t in otys]
    ip = (C.c_void_p * 32)(*[a.ctypes.data for a in ins])
    op = (C.c_void_p * 32)(*[a.ctypes.data for a in outs])
    ex, er = C.c_uint64(), C.c_uint32()
    rc = lib(path).oracle_run(k, ip, op, 0, n, threads, C.byref(ex), C.byref(er))
    assert rc == 0
    return outs, er.value


def literals(name, path=PATH):
    """(text, dtype, bits) of every literal operand of kernel `name`, in
    statement order, with the bits the kernel computes with."""
    L = lib(path)
    k = L.oracle_kernel_find(name.encode())
    if k < 0:
        raise KeyError(name)
    out, buf, ty, bits = [], C.create_string_buffer(256), C.c_int(), C.c_uint64()
    i = 0
    while L.oracle_kernel_literal(k, i, buf, 256, C.byref(ty), C.byref(bits)) == 0:
        out.append((buf.value.decode(), DTYPES[ty.value], bits.value))
        i += 1
    return out


def collatz_direct(x, threads=8):
    """Collatz step counts by direct iteration of the definition."""
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.zeros(x.size, np.uint32)
    rc = lib().oracle_collatz_direct(x.ctypes.data, out.ctypes.data, x.size, threads)
    assert rc == 0, "a trajectory left 64 bits"
    return out


def kernel_names(path=PATH):
    L = lib(path)
    return [L.oracle_kernel_name(k).decode() for k in range(L.oracle_kernel_count())]


def parse(kind, text):
    v = {"f32": C.c_float, "f64": C.c_double, "u64": C.c_uint64, "i32": C.c_int32}[kind]()
    rc = getattr(lib(), f"oracle_parse_{kind}")(text.encode(), C.byref(v))
    return None if rc else v.value


def smallpt9_direct(sid, job, threads=8):
    """The direct C++ path tracer (not the IR) on host columns."""
    sid = np.ascontiguousarray(sid, np.uint32)
    job = np.ascontiguousarray(job, np.uint32)
    n = len(sid)
    outs = [np.zeros(n, np.float32) for _ in range(3)]
    lib().oracle_smallpt9_direct(sid.ctypes.data, job.ctypes.data, outs[0].ctypes.data,
                                 outs[1].ctypes.data, outs[2].ctypes.data, 0, n, threads)
    return outs


def trace(name, inputs, thread, max_steps):
    """Step trace of one thread (oracle): a list of (label set, error bits,
    {register name: bits}) after every block."""
    k, itys, _ = io(name)
    L = lib()
    ins = [np.ascontiguousarray(a, dtype=t) for a, t in zip(inputs, itys)]
    ni, no, nr = C.c_int(), C.c_int(), C.c_int()
    it, ot = (C.c_int * 32)(), (C.c_int * 32)()
    L.oracle_kernel_io(k, C.byref(ni), C.byref(no), it, ot, C.byref(nr))
    names = []
    buf = C.create_string_buffer(256)
    for r in range(nr.value):
        assert L.oracle_register_name(k, r, buf, 256) == 0
        names.append(buf.value.decode())
    first = np.zeros(max_steps, np.int32)
    regs = np.zeros((max_steps, nr.value), np.uint64)
    errs = np.zeros(max_steps, np.uint32)
    steps = C.c_uint32()
    ip = (C.c_void_p * 32)(*[a.ctypes.data for a in ins])
    assert L.oracle_trace(k, ip, thread, max_steps, first.ctypes.data, regs.ctypes.data,
                          errs.ctypes.data, C.byref(steps)) == 0
    out = []
    for s in range(steps.value):
        assert L.oracle_block_labels(k, int(first[s]), buf, 256) == 0
        out.append((set(buf.value.decode().split()), int(errs[s]),
                    dict(zip(names, (int(v) for v in regs[s])))))
    return out


def work(name, inputs, threads=8):
    """(thread steps, ops, flops) of every thread of `inputs` (oracle)."""
    k, itys, _ = io(name)
    ins = [np.ascontiguousarray(a, dtype=t) for a, t in zip(inputs, itys)]
    ip = (C.c_void_p * 32)(*[a.ctypes.data for a in ins])
    st, op, fl = C.c_uint64(), C.c_uint64(), C.c_uint64()
    assert lib().oracle_work(k, ip, 0, len(ins[0]), threads, C.byref(st), C.byref(op),
                             C.byref(fl)) == 0
    return st.value, op.value, fl.value
