"""Build recipe: compile the gfx950 kernel library and the CPU oracle in-tree.

    python -m loom_amd.build            # library + oracle
    python -m loom_amd.build --lib      # library only

The library is `loom_amd/_build/libloom_amd.so` (C ABI in include/loom_amd.h).
Every HIP source is compiled for gfx950 with contraction off (the op
contract of the VM is IEEE round-to-nearest without fused multiply-add unless
the IR asks for fma) and without packed FP32 instructions.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
BUILD = PKG / "_build"
LIB = BUILD / "libloom_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++20",
    f"--offload-arch={ARCH}",
    "-O3",
    "-ffp-contract=off",
    "-fPIC",
    "-fconstexpr-steps=200000000",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep the stores of different switch cases apart: merging them into one
    # store with a computed register index would move a kernel's register
    # file from VGPRs to scratch memory
    "-mllvm", "-simplifycfg-sink-common=false",
    f"-I{ROOT / 'include' / 'loom_vm'}",
    f"-I{ROOT / 'include'}",
]


def sources() -> list[Path]:
    csrc = PKG / "csrc"
    return [csrc / "runtime.cpp", csrc / "serve.cpp", csrc / "stream.cpp"] + sorted(csrc.glob("kernels_*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{r.stderr[-8000:]}")
    return obj


def build_lib(jobs: int | None = None) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = BUILD / "libloom_amd.so.tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr[-8000:]}")
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    r = subprocess.run(["make", "-C", str(ORACLE_DIR), "-B", "-j4"], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stderr[-8000:]}")
    return ORACLE_LIB


FACADE_SRC = ROOT / "tests" / "cpp" / "facade_test.cpp"
FACADE_BIN = ROOT / "tests" / "cpp" / "_build" / "facade_test"


def build_facade_test() -> Path:
    """Host program that drives the library through the C++ facade
    (include/loom_vm/vm/hip.hpp); host compiler only, links the library."""
    FACADE_BIN.parent.mkdir(parents=True, exist_ok=True)
    tmp = FACADE_BIN.with_suffix(".tmp")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++20", "-O2", "-Wall",
           f"-I{ROOT / 'include' / 'loom_vm'}", f"-I{ROOT / 'include'}",
           str(FACADE_SRC), "-o", str(tmp), f"-L{BUILD}", "-lloom_amd",
           f"-Wl,-rpath,{BUILD}", "-Wl,-rpath,$ORIGIN/../../../loom_amd/_build"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade test build failed\n{r.stderr[-8000:]}")
    os.replace(tmp, FACADE_BIN)
    return FACADE_BIN


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_lib(a.j), flush=True)
        print("built", build_facade_test(), flush=True)
    if not a.lib:
        print("built", build_oracle(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
