"""The Shop's host-side coordination under ThreadSanitizer (CPU only): the
facade's basic_shop / basic_context (include/loom_vm/vm/hip.hpp, the
reference's mp::Shop protocol) driven by worker threads and an exchanging
main thread over a CPU stand-in stream (tests/cpp/shop_tsan.cpp).  A data
race or a lost or duplicated output fails the test."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_shop_protocol_under_tsan(tmp_path):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    exe = tmp_path / "shop_tsan"
    cmd = [cxx, "-std=c++20", "-O1", "-g", "-fsanitize=thread",
           f"-I{ROOT / 'include' / 'loom_vm'}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "shop_tsan.cpp"), "-o", str(exe), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "shop_tsan: ok" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr
