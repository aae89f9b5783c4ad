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


@pytest.mark.gpu
def test_progressive_render_equals_single_launch(gpu, tmp_path):
    """The progressive renderer (four launches of two samples per pixel)
    runs the same sample threads as one launch of the whole frame: every
    intermediate image is the mean of the samples so far, and the last one
    equals the single launch and the oracle bit for bit."""
    import numpy as np
    from loom_amd import image, workloads
    from tests import oracle_lib as O
    w, h, spp, passes = 40, 24, 8, 4
    print("[gpu] progressive render", flush=True)
    seen = []
    last = image.render_progressive(w, h, spp, passes,
                                    on_pass=lambda k, img, st: seen.append((k, img.copy(),
                                                                            st.errors)))
    assert [k for k, _, _ in seen] == list(range(passes))
    assert all(e == 0 for _, _, e in seen)
    want = workloads.pt_image(*O.run("smallpt9_f32", workloads.pt_inputs(w, h, spp))[0], w, h,
                              spp)
    assert np.array_equal(last, want)
    assert np.array_equal(image.render(w, h, spp), want)
    # the first pass is the mean of the first two samples of each pixel
    (ox, oy, oz), _ = O.run("smallpt9_f32", workloads.pt_inputs(w, h, spp))
    first = workloads.pt_image(*(c.reshape(w * h, spp)[:, :2].ravel() for c in (ox, oy, oz)),
                               w, h, 2)
    assert np.array_equal(seen[0][1], first)
    image.main([str(tmp_path / "p.ppm"), "--w", "16", "--h", "8", "--spp", "4",
                "--progressive", "2"])
    assert (tmp_path / "p.ppm").stat().st_size == len(b"P6\n16 8\n255\n") + 16 * 8 * 3
