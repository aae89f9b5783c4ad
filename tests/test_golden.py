"""The committed golden fixtures (tests/golden/, written by make_golden.py
from computations independent of the oracle's interpreter) against the
workload generator and the oracle."""
import hashlib
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from loom_amd import workloads
from tests import oracle_lib as O

G = Path(__file__).resolve().parent / "golden"


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    return json.loads((G / name).read_text())


def test_fib_workload_and_oracle():
    f = load("fib_u64_c1.json")
    n = workloads.fib_inputs(f["threads"])
    assert sha(n) == f["inputs_sha256"]  # generator is stable across machines
    (c,), err = O.run("fib_u64", [n])
    assert err == 0
    assert sha(c) == f["outputs_sha256"]


def test_collatz_workload_and_oracle_sample():
    f = load("collatz_u64_c2.json")
    x = workloads.collatz_inputs(f["threads"])
    assert sha(x) == f["inputs_sha256"]
    rows = np.array(f["sample"], dtype=np.uint64)
    assert np.array_equal(x[rows[:, 0].astype(np.int64)], rows[:, 1])
    (s,), err = O.run("collatz_u64", [rows[:, 1]])
    assert err == 0
    assert np.array_equal(s.astype(np.uint64), rows[:, 2])


def test_smallpt9_frame_oracle_interpreter():
    """The IR interpreter reproduces the fixture frame (made by the direct
    C++ restatement) bit for bit."""
    z = np.load(G / "smallpt9_32x24x8.npz")
    w, h, spp = int(z["w"]), int(z["h"]), int(z["spp"])
    ins = workloads.pt_inputs(w, h, spp)
    assert int(ins[1][0]) == int(z["job"])
    out, err = O.run("smallpt9_f32", ins)
    assert err == 0
    for got, key in zip(out, ("lx", "ly", "lz")):
        assert np.array_equal(got.view(np.uint32), z[key])
    img = workloads.pt_image(*(z[k].view(np.float32) for k in ("lx", "ly", "lz")), w, h, spp)
    assert np.array_equal(img, z["image"])


def test_literals_table():
    """The oracle's parser (the C++ restatement the product shares) against
    the fixture's values from the Python restatement of the reference
    parser, bit for bit, in both precisions."""
    f = load("literals.json")
    for row in f["literals"]:
        t = row["text"]
        f32 = O.parse("f32", t)
        f64 = O.parse("f64", t)
        assert struct.unpack("<I", struct.pack("<f", f32))[0] == row["s2f_f32_bits"], t
        assert struct.unpack("<Q", struct.pack("<d", f64))[0] == row["s2f_f64_bits"], t
        # the reference parser stays within a few ulp of correct rounding
        assert abs(row["s2f_f32_bits"] - row["rn_f32_bits"]) <= 4, t
        assert abs(row["s2f_f64_bits"] - row["rn_f64_bits"]) <= 8, t


def test_literals_table_is_not_correct_rounding():
    """The fixture separates the reference's parser from a correctly rounding
    one: a parser that rounded correctly would fail the table above on
    these texts."""
    f = load("literals.json")
    inexact = {r["text"] for r in f["literals"]
               if r["s2f_f32_bits"] != r["rn_f32_bits"] or r["s2f_f64_bits"] != r["rn_f64_bits"]}
    assert {"0.093073", "0.5135", "0.999"} <= inexact


@pytest.mark.parametrize("text,exact", [("0.5", True), ("0.25", True), ("12", True),
                                        ("16777216", True), ("1e5", True)])
def test_literals_exact_where_representable(text, exact):
    row = next(r for r in load("literals.json")["literals"] if r["text"] == text)
    assert (row["s2f_f32_bits"] == row["rn_f32_bits"]) == exact


def test_ppm_of_fixture_frame(tmp_path):
    """The image writer on the fixture frame: header, size, orientation (the
    ceiling light is at the top of the file), deterministic bytes."""
    from loom_amd import image
    z = np.load(G / "smallpt9_32x24x8.npz")
    img = z["image"]
    path = tmp_path / "f.ppm"
    image.write_ppm(path, img)
    raw = path.read_bytes()
    head = b"P6\n32 24\n255\n"
    assert raw.startswith(head) and len(raw) == len(head) + 32 * 24 * 3
    px = np.frombuffer(raw[len(head):], np.uint8).reshape(24, 32, 3)
    assert px[:3].mean() > px[-3:].mean()  # light near the top row of the file
    assert np.array_equal(px, image.to_bytes(img[::-1]))


def test_work_totals_fib_full_config():
    """The oracle's count of blocks, ops and flops over the whole fib config
    equals the closed form of the fixture."""
    f = load("work_totals.json")["fib_u64_c1"]
    assert f["threads"] == 1 << 20
    got = O.work("fib_u64", [workloads.fib_inputs(1 << 20)])
    assert got == (f["steps"], f["ops"], f["flops"])


def test_work_totals_collatz_closed_form_on_prefix():
    """On a prefix of the Collatz config the oracle's count equals the closed
    form (steps 2 + 3c, ops 5 + 7c + odd steps) of brute-force step counts."""
    from tests.golden.make_golden import collatz_odd_steps, collatz_steps
    x = workloads.collatz_inputs(1 << 24)[: 1 << 14]
    c = collatz_steps(x).astype(np.uint64)
    odd = collatz_odd_steps(x)
    got = O.work("collatz_u64", [x])
    assert got == (int((2 + 3 * c).sum()), int((5 + 7 * c + odd).sum()), 0)


def test_work_totals_path_tracer_prefix():
    f = load("work_totals.json")["smallpt9_c3_first_65536"]
    sid, job = workloads.pt_inputs(1024, 768, 64)
    got = O.work("smallpt9_f32", [sid[:65536], job[:65536]])
    assert got == (f["steps"], f["ops"], f["flops"])
    # every sample runs START, then at least one intersection
    assert f["steps"] >= 2 * 65536 and f["flops"] > 0


def test_full_frame_fixture_against_the_ir():
    """The whole-frame fixture (direct C++ restatement) at its sampled rows
    equals the oracle's block interpreter running the kernel's IR."""
    f = json.loads((G / "smallpt9_1024x768x64.json").read_text())
    sid, job = workloads.pt_inputs(f["w"], f["h"], f["spp"])
    assert hashlib.sha256(sid.tobytes()).hexdigest() == f["inputs_sha256"]
    assert int(job[0]) == f["job"] and len(sid) == f["threads"]
    rows = np.array([r[0] for r in f["sample"]], np.int64)
    got, err = O.run("smallpt9_f32", [sid[rows], job[rows]])
    assert err == 0
    for c in range(3):
        want = np.array([r[1 + c] for r in f["sample"]], np.uint32)
        assert np.array_equal(got[c].view(np.uint32), want)
