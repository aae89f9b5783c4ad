"""Writes the golden fixtures of this directory.

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py work     # work_totals.json only
    python tests/golden/make_golden.py frame    # smallpt9_1024x768x64.json only
    python tests/golden/make_golden.py literals # literals.json only

The reference holds no golden vectors for this path and cannot be built
here, so the fixtures come from computations independent of the oracle's
block interpreter and of the GPU backend:

* fib_u64_c1.json — the fib workload (2^20 seeded inputs in [0, 90]); the
  expected outputs come from the Fibonacci recurrence in Python integers.
* collatz_u64_c2.json — the Collatz workload (2^24 seeded inputs in
  [1, 2^32)); the expected step counts come from a vectorised brute-force
  iteration in numpy with an overflow check.
* smallpt9_32x24x8.npz — a 32x24 frame at 8 spp of the 9-sphere path
  tracer from the direct C++ restatement (not the IR): per-sample radiance
  bits and the per-pixel image.
* smallpt9_1024x768x64.json — the whole path-tracing workload frame
  (50,331,648 sample threads) from the direct C++ restatement: SHA-256 of
  each radiance channel's bits and a sample of (sample index, bits) rows.
* literals.json — literal strings with the value of the reference's literal
  parser from a Python restatement in exact rational arithmetic
  (tests/literal_ref.py, not the C++ restatement the oracle and the product
  share) and the correctly rounded float32 and float64 values.
* work_totals.json — the algorithmic work of whole configurations (thread
  block steps, ops = #Compute + 1 per block run, floating-point ops with
  fma = 2): fib and Collatz in closed form from the IR text and the
  brute-force step counts (Collatz: odd steps counted separately), the path
  tracer (prefixes of the 1024x768x64 frame and the whole frame) by the
  oracle's block interpreter.

Each file stores SHA-256 digests of the full input and output columns and a
sample of (index, input, output) rows.
"""
from __future__ import annotations

import hashlib
import json
import struct
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from loom_amd import workloads  # noqa: E402

SAMPLE = 1024


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def sample_rows(n: int, seed: int) -> np.ndarray:
    idx = workloads.uniform_int(seed, SAMPLE, 0, n - 1).astype(np.int64)
    return np.unique(np.concatenate([idx, [0, n - 1]]))


def fib_fixture():
    n = workloads.fib_inputs(1 << 20)
    table = [0, 1]
    while len(table) < 91:
        table.append(table[-1] + table[-2])
    assert table[90] < 1 << 64
    c = np.array([table[int(v)] for v in range(91)], dtype=np.uint64)[n.astype(np.int64)]
    rows = sample_rows(len(n), 101)
    return {"kernel": "fib_u64", "threads": len(n), "inputs_sha256": sha(n),
            "outputs_sha256": sha(c),
            "sample": [[int(i), int(n[i]), int(c[i])] for i in rows],
            "how": "Fibonacci recurrence in Python integers"}


def collatz_steps(x: np.ndarray) -> np.ndarray:
    """Step counts of the Collatz map until 1, vectorised, with an overflow
    check on 3x+1."""
    x = x.astype(np.uint64).copy()
    steps = np.zeros(len(x), np.uint32)
    idx = np.flatnonzero(x > 1)
    cur = x[idx]
    limit = np.uint64((2**64 - 1 - 1) // 3)
    while len(idx):
        odd = (cur & np.uint64(1)).astype(bool)
        if (cur[odd] > limit).any():
            raise OverflowError("3x+1 overflows u64")
        cur = np.where(odd, cur * np.uint64(3) + np.uint64(1), cur >> np.uint64(1))
        steps[idx] += 1
        keep = cur > 1
        idx, cur = idx[keep], cur[keep]
    return steps


def collatz_odd_steps(x: np.ndarray) -> np.ndarray:
    """Number of 3x+1 steps of each start value (vectorised)."""
    x = x.astype(np.uint64).copy()
    odd_n = np.zeros(len(x), np.uint64)
    idx = np.flatnonzero(x > 1)
    cur = x[idx]
    while len(idx):
        odd = (cur & np.uint64(1)).astype(bool)
        odd_n[idx[odd]] += 1
        cur = np.where(odd, cur * np.uint64(3) + np.uint64(1), cur >> np.uint64(1))
        keep = cur > 1
        idx, cur = idx[keep], cur[keep]
    return odd_n


def work_fixture(full_frame: bool = True):
    """Work totals; see the module docstring."""
    from tests import oracle_lib as O
    out = {}
    # fib (programs.hpp): START 3 ops; n > 1: work 4 ops, then n-1 loop trips of 6
    n = workloads.fib_inputs(1 << 20).astype(np.int64)
    big = n > 1
    out["fib_u64_c1"] = {
        "threads": int(len(n)),
        "steps": int(np.where(big, n + 1, 1).sum()),
        "ops": int(np.where(big, 7 + 6 * (n - 1), 3).sum()),
        "flops": 0,
        "how": "closed form: steps 1 or n+1, ops 3 or 7+6(n-1)"}
    # Collatz: START 3 + done 2, per step: step 2 + next 3 + (even 2 | odd3 3)
    x = workloads.collatz_inputs(1 << 24)
    cnt = collatz_steps(x).astype(np.uint64)
    odd = collatz_odd_steps(x)
    out["collatz_u64_c2"] = {
        "threads": int(len(x)),
        "steps": int((2 + 3 * cnt).sum(dtype=np.uint64)),
        "ops": int((5 + 7 * cnt + odd).sum(dtype=np.uint64)),
        "flops": 0,
        "how": "closed form from brute-force step counts: steps 2+3c, ops 5+7c+odd"}
    sid, job = workloads.pt_inputs(1024, 768, 64)
    sizes = [1 << 16, 1 << 20] + ([len(sid)] if full_frame else [])
    for m in sizes:
        st, op, fl = O.work("smallpt9_f32", [sid[:m], job[:m]])
        key = "smallpt9_c3" if m == len(sid) else f"smallpt9_c3_first_{m}"
        out[key] = {"threads": m, "steps": st, "ops": op, "flops": fl,
                    "how": "oracle block interpreter, " + (
                        "the whole 1024x768x64 frame" if m == len(sid) else
                        "first samples of the 1024x768x64 frame")}
        print(key, out[key], flush=True)
    return out


def collatz_fixture():
    x = workloads.collatz_inputs(1 << 24)
    s = collatz_steps(x)
    rows = sample_rows(len(x), 202)
    return {"kernel": "collatz_u64", "threads": len(x), "inputs_sha256": sha(x),
            "outputs_sha256": sha(s), "max_steps": int(s.max()), "sum_steps": int(s.sum(dtype=np.uint64)),
            "sample": [[int(i), int(x[i]), int(s[i])] for i in rows],
            "how": "vectorised brute-force Collatz iteration (numpy), overflow-checked"}


def smallpt_fixture():
    from tests import oracle_lib as O
    w, h, spp = 32, 24, 8
    sid, job = workloads.pt_inputs(w, h, spp)
    lx, ly, lz = O.smallpt9_direct(sid, job)
    img = workloads.pt_image(lx, ly, lz, w, h, spp)
    np.savez_compressed(HERE / "smallpt9_32x24x8.npz",
                        w=w, h=h, spp=spp, job=np.uint32(job[0]),
                        lx=lx.view(np.uint32), ly=ly.view(np.uint32), lz=lz.view(np.uint32),
                        image=img)


def frame_fixture():
    """The 1024x768x64 frame by the direct C++ path tracer (not the IR)."""
    from tests import oracle_lib as O
    w, h, spp = 1024, 768, 64
    sid, job = workloads.pt_inputs(w, h, spp)
    cols = O.smallpt9_direct(sid, job, threads=8)
    rows = sample_rows(len(sid), 303)
    return {"kernel": "smallpt9_f32", "w": w, "h": h, "spp": spp, "threads": len(sid),
            "job": int(job[0]), "inputs_sha256": sha(sid),
            "outputs_sha256": [sha(c.view(np.uint32)) for c in cols],
            "sample": [[int(i)] + [int(c.view(np.uint32)[i]) for c in cols] for i in rows],
            "how": "direct C++ restatement of the path tracer (oracle smallpt9_direct), "
                   "radiance bits per channel"}


LITERALS = ["0", "1", "0.5", "0.1", "0.25", "2.5", "1e-12", "1e-6", "1e5", "1e30", "-0.1",
            "6.283185307179586", "3.14159265", "0.093073", "0.999", "81.6", "681.33",
            "0.042612", "0.5135", "295.6", "1.5", "0.75", "12", "16777216", "16777217",
            "0.0001", "2.718281828", "100000.5", "1e-3", "0.333333333", "7.5e-4", "1e1"]


def literal_fixture():
    """Literal texts with two values each, neither from the product's or the
    oracle's parser: s2f_* from the Python restatement of the reference's
    parser (tests/literal_ref.py, exact rationals rounded at every step) and
    rn_* the correctly rounded value."""
    from tests import literal_ref as R
    out = []
    for t in LITERALS:
        row = {"text": t}
        for kind in ("f32", "f64"):
            row[f"s2f_{kind}_bits"] = R.bits(R.s2f(t, kind), kind)
            row[f"rn_{kind}_bits"] = R.bits(R.correctly_rounded(t, kind), kind)
        out.append(row)
    return {"literals": out,
            "how": "s2f_*: the reference parser's value (string.hpp:312-352) from the Python "
                   "restatement tests/literal_ref.py; rn_*: correctly rounded"}


def main():
    if sys.argv[1:] == ["frame"]:
        (HERE / "smallpt9_1024x768x64.json").write_text(json.dumps(frame_fixture(), indent=0) + "\n")
        print("frame", flush=True)
        return
    if sys.argv[1:] == ["literals"]:
        (HERE / "literals.json").write_text(json.dumps(literal_fixture(), indent=0) + "\n")
        print("literals", flush=True)
        return
    if sys.argv[1:] == ["work"]:
        (HERE / "work_totals.json").write_text(json.dumps(work_fixture(), indent=1) + "\n")
        print("work totals", flush=True)
        return
    (HERE / "fib_u64_c1.json").write_text(json.dumps(fib_fixture(), indent=0) + "\n")
    print("fib", flush=True)
    (HERE / "collatz_u64_c2.json").write_text(json.dumps(collatz_fixture(), indent=0) + "\n")
    print("collatz", flush=True)
    smallpt_fixture()
    print("smallpt9", flush=True)
    (HERE / "smallpt9_1024x768x64.json").write_text(json.dumps(frame_fixture(), indent=0) + "\n")
    print("frame", flush=True)
    (HERE / "literals.json").write_text(json.dumps(literal_fixture(), indent=0) + "\n")
    print("literals", flush=True)
    (HERE / "work_totals.json").write_text(json.dumps(work_fixture(), indent=1) + "\n")
    print("work totals", flush=True)


if __name__ == "__main__":
    main()
