"""The C++ facade (include/loom_vm/vm/hip.hpp) driven by a host program the
way the reference's VM tests drive its SIMD backends: compact and hybrid
functions and the stream protocol (reference src/tests/vm.cpp:40-260,
src/tests/stream_wrapper.hpp:33-80).  The program checks closed forms and
brute force itself and exits non-zero on a mismatch."""
import subprocess

import pytest

from loom_amd import build


@pytest.mark.gpu
def test_facade_program():
    exe = build.FACADE_BIN
    if not exe.exists():
        pytest.fail(f"{exe} is not built (python -m loom_amd.build)")
    print("running", exe, flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "facade_test: ok" in r.stdout


def test_facade_builds(tmp_path, monkeypatch):
    """The facade compiles with the host compiler against the C ABI."""
    monkeypatch.setattr(build, "FACADE_BIN", tmp_path / "facade_test")
    if not build.LIB.exists():
        pytest.skip("library not built")
    assert build.build_facade_test().exists()


@pytest.mark.gpu
def test_render_and_write_ppm(gpu, tmp_path):
    """End to end: a frame rendered by the hybrid form equals the oracle's
    image, and the PPM writer produces the file."""
    import numpy as np
    from loom_amd import image, workloads
    from tests import oracle_lib as O
    w, h, spp = 48, 32, 4
    print("[gpu] render ppm", flush=True)
    img = image.render(w, h, spp)
    want = workloads.pt_image(*O.run("smallpt9_f32", workloads.pt_inputs(w, h, spp))[0], w, h, spp)
    assert np.array_equal(img, want)
    image.write_ppm(tmp_path / "frame.ppm", img)
    assert (tmp_path / "frame.ppm").stat().st_size == len(b"P6\n48 32\n255\n") + w * h * 3
