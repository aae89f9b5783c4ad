"""The reference's own test programs (src/tests/programs.hpp) on the GPU
through the C ABI, in every form, bit-exact against the oracle interpreter
running the same programs and against the reference's own validators — the
checks its CUDA backend test makes (src/tests/vm.cu:58-86: function and
stream forms against fn::val_*), over every instantiation its VM suites use."""
import numpy as np
import pytest

import loom_amd
from loom_amd import abi
from tests import oracle_lib as O
from tests import refprograms as P
from tests.test_refkernels import REF_LIB, load_ref_kernels

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (REF_LIB.exists() and O.REF_PATH.exists()),
                                 reason="reference kernels not built (no reference checkout)")]

# 2^16 threads: with one wave per CU the hybrid form holds 16384 lanes, so
# every lane takes about four threads from the stream in turn
N = 1 << 16
FORMS = [
    ("compact_sm", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR)),
    ("compact_topo", abi.MODE_COMPACT, dict(reconv=abi.RECONV_TOPO)),
    ("compact_sweep", abi.MODE_COMPACT, dict(reconv=abi.RECONV_SWEEP)),
    ("hybrid_sweep", abi.MODE_HYBRID, {}),
    ("hybrid_stream_1wpc", abi.MODE_HYBRID, dict(waves_per_cu=1, live_threshold=0)),
    ("hybrid_sm_regroup", abi.MODE_HYBRID, dict(reconv=abi.RECONV_SUCCESSOR_MAJOR,
                                                regroup=True, live_threshold=32)),
    ("hybrid_workgroup_regroup", abi.MODE_HYBRID, dict(regroup=abi.REGROUP_WORKGROUP,
                                                       reconv=abi.RECONV_SWEEP,
                                                       regroup_phase=16)),
]


@pytest.mark.parametrize("name", P.NAMES)
def test_reference_program_parity(gpu, name):
    load_ref_kernels()
    cols = P.inputs(name, N, seed=1)
    (want,), err = O.run(name, cols, threads=16, path=O.REF_PATH)
    assert err == 0
    expect = P.validate(name, cols)
    if expect is not None and P.PROGRAMS[P.split(name)[0]][1] == "fuzzy":
        assert P.ulp_mismatches(want, expect) <= 0.01 * N, name
    elif expect is not None:
        assert np.array_equal(want, expect), name
    k = loom_amd.kernel(name)
    for fname, mode, params in FORMS:
        print(f"[gpu] {name} {fname} n={N}", flush=True)
        r = loom_amd.run(k, mode, cols, params=loom_amd.make_params(mode=mode, **params))
        assert r.errors == 0, (name, fname)
        (got,) = r.outputs
        assert got.dtype == want.dtype
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (name, fname)


def test_reference_rng_macros(gpu):
    """The reference's random-number IR macros on the GPU, every form,
    bit-exact against the oracle and the reference's C++ functions."""
    load_ref_kernels()
    cols = P.rng_inputs(N, seed=1)
    want, err = O.run(P.RNG, cols, threads=16, path=O.REF_PATH)
    assert err == 0
    ref = P.rng_reference(*cols)
    for w, r in zip(want, ref):
        assert np.array_equal(w.view(np.uint32), r.view(np.uint32))
    k = loom_amd.kernel(P.RNG)
    for fname, mode, params in FORMS:
        print(f"[gpu] {P.RNG} {fname} n={N}", flush=True)
        r = loom_amd.run(k, mode, cols, params=loom_amd.make_params(mode=mode, **params))
        assert r.errors == 0
        for g, w in zip(r.outputs, want):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), fname


@pytest.mark.parametrize("name", P.JMP4)
def test_reference_jmp4_past_the_table(gpu, name):
    """Jump indexes over the whole index range on the GPU (indexes past the
    table halt), every form, against the oracle and the reference's
    validator."""
    load_ref_kernels()
    cols = P.jmp4_wide_inputs(name, N, seed=1)
    (want,), err = O.run(name, cols, threads=16, path=O.REF_PATH)
    assert err == 0
    assert np.array_equal(want, P.validate(name, cols))
    k = loom_amd.kernel(name)
    for fname, mode, params in FORMS:
        print(f"[gpu] {name} wide {fname} n={N}", flush=True)
        r = loom_amd.run(k, mode, cols, params=loom_amd.make_params(mode=mode, **params))
        assert r.errors == 0
        assert np.array_equal(r.outputs[0], want), fname


def test_reference_fib_at_full_size(gpu):
    """The fib workload at its full size (2^20 threads, n in [0, 90]) through
    the reference's own fib_test program, compact and hybrid, against the
    reference's own validator val_fib; this repository's fib_u64 kernel
    gives the same outputs on the same inputs."""
    from loom_amd import workloads
    load_ref_kernels()
    n = workloads.fib_inputs()
    want = P.validate("ref_fib_test_u64", [n])
    k = loom_amd.kernel("ref_fib_test_u64")
    for fname, mode, params in FORMS:
        print(f"[gpu] ref_fib_test_u64 2^20 {fname}", flush=True)
        r = loom_amd.run(k, mode, [n], params=loom_amd.make_params(mode=mode, **params))
        assert r.errors == 0
        assert np.array_equal(r.outputs[0], want), fname
    own = loom_amd.run("fib_u64", abi.MODE_HYBRID, [n])
    assert np.array_equal(own.outputs[0], want)


def test_reference_programs_through_the_cpp_facade(gpu):
    """oracle/_ref/ref_facade_test: the reference's own test programs driven
    through the C++ facade with the reference's call shapes (a function per
    program, the stream protocol for factorial, as its CUDA backend test
    does), checked against its validators; exits non-zero on a mismatch."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "ref_facade_test"
    if not exe.exists():
        pytest.skip("reference facade test not built (no reference checkout)")
    print("[gpu] running", exe, flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ref_facade_test: ok" in r.stdout


NARROW = [n for n in P.NAMES if n.endswith(("_u8", "_i8", "_u16", "_i16"))]


@pytest.mark.parametrize("name", NARROW)
def test_reference_program_stream(gpu, name):
    """The device-resident stream over the reference's 8- and 16-bit
    programs: each register's column holds its 32-bit cell, and every output
    equals the oracle's."""
    from loom_amd.stream import Stream
    load_ref_kernels()
    cols = P.inputs(name, 5000, seed=3)
    (want,), err = O.run(name, cols, threads=16, path=O.REF_PATH)
    assert err == 0
    print(f"[gpu] stream {name}", flush=True)
    s = Stream(name, 256, live_threshold=16)
    (got,) = s.run_all(cols)
    s.close()
    assert got.dtype == want.dtype
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), name
