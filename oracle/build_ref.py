"""ORACLE / TEST INFRASTRUCTURE: build the reference's own path-tracing
kernels and test programs (oracle/ref_kernels.hpp) against this repository's
front-end, from the reference checkout where it lies, into oracle/_ref/:

* liboracle_ref.so — the oracle interpreter (oracle.cpp) over those kernels,
  the reference's own validators of the test programs (ref_validators.cpp),
  and the C++ side of its path-tracing and vector macro checks
  (ref_macro_checks.cpp);
* libloom_refkernels.so — the same kernels lowered for gfx950 by the
  product's backend; loading it (after the product library) registers them
  with the product library, where the C ABI finds them by name;
* ref_facade_test — a host program that drives the reference's test
  programs through the C++ facade (ref_facade_test.cpp).

Only tests load these.  Nothing is built when the reference is absent (the
GPU box has none; the libraries built here travel with the tree).

    python oracle/build_ref.py [--ref /root/reference] [--oracle-only]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
OUT = HERE / "_ref"
ORACLE_LIB = OUT / "liboracle_ref.so"
KERNEL_LIB = OUT / "libloom_refkernels.so"
FACADE_BIN = OUT / "ref_facade_test"
CLANG = os.environ.get("LOOM_HOST_CLANG", "/opt/rocm/llvm/bin/clang++")
SOURCES = ["ref_pt_const.hip", "ref_pt2_const.hip", "ref_pt2_const_1024.hip",
           "ref_programs_a.hip", "ref_programs_b.hip", "ref_programs_c.hip", "ref_macros.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:3])} ... failed\n{r.stderr[-6000:]}")


def build(ref: Path, before_link=None, oracle_only: bool = False) -> list[Path]:
    """Build both libraries; `before_link` (if given) is called after the
    objects are compiled and before the kernel library is linked against
    libloom_amd.so, so a caller can compile this concurrently with that
    library and make the link wait for it."""
    sys.path.insert(0, str(ROOT))
    from loom_amd import build as B
    src = ref / "src"
    # static_math: the constexpr math library the reference vendors for its
    # common/math.hpp (loom::fac, used by the validators)
    ext = ref / "ext" / "include"
    if not (src / "smallpt" / "vm" / "pt.hpp").exists():
        return []
    OUT.mkdir(parents=True, exist_ok=True)

    def oracle():
        # the C++ side of the macro checks is compiled as the reference
        # compiles its own (clang, contraction of a*b+c within an expression,
        # a CPU with fused multiply-add: CMakeLists.txt:64-77), the
        # interpreter and the validators with the op contract's flags
        checks = OUT / "ref_macro_checks.o"
        _run([CLANG, "-O2", "-std=c++20", "-fPIC", "-march=x86-64-v3", "-ffp-contract=on",
              f"-I{HERE}", f"-I{src}", f"-I{ext}", "-c", str(HERE / "ref_macro_checks.cpp"),
              "-o", str(checks)])
        tmp = ORACLE_LIB.with_suffix(".tmp")
        _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++20", "-fPIC", "-ffp-contract=off",
              "-fno-fast-math", "-DLOOM_ORACLE_REF", f"-I{ROOT / 'include' / 'loom_vm'}",
              f"-I{ROOT / 'include'}", f"-I{HERE}", f"-I{src}", f"-I{ext}", "-shared", "-o",
              str(tmp), str(HERE / "oracle.cpp"), str(HERE / "ref_validators.cpp"), str(checks),
              "-lpthread"])
        os.replace(tmp, ORACLE_LIB)
        checks.unlink()
        return ORACLE_LIB

    def obj(name):
        o = OUT / (Path(name).stem + ".o")
        _run([B.HIPCC, *B.COMMON, f"-I{ROOT / 'loom_amd' / 'csrc'}", f"-I{HERE}", f"-I{src}",
              f"-I{ext}", "-x", "hip", "-c", str(HERE / name), "-o", str(o)])
        return o

    if oracle_only:
        return [oracle()]
    with cf.ThreadPoolExecutor(min(len(SOURCES) + 1, 8)) as ex:
        fo = ex.submit(oracle)
        objs = list(ex.map(obj, SOURCES))
        fo.result()
    if before_link is not None:
        before_link()
    tmp = KERNEL_LIB.with_suffix(".tmp")
    _run([B.HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          f"-L{B.BUILD}", "-lloom_amd", f"-Wl,-rpath,{B.BUILD}",
          "-Wl,-rpath,$ORIGIN/../../loom_amd/_build"])
    os.replace(tmp, KERNEL_LIB)
    for o in objs:
        o.unlink()
    # linked with --no-as-needed: nothing references the kernel library by
    # symbol, its load registers the kernels the facade looks up
    tmp = FACADE_BIN.with_suffix(".tmp")
    _run([os.environ.get("CXX", "g++"), "-std=c++20", "-O2", "-ffp-contract=off",
          f"-I{ROOT / 'include' / 'loom_vm'}", f"-I{ROOT / 'include'}", f"-I{src}", f"-I{ext}",
          str(HERE / "ref_facade_test.cpp"), "-o", str(tmp), f"-L{B.BUILD}", f"-L{OUT}",
          "-Wl,--no-as-needed", "-lloom_refkernels", "-lloom_amd", "-Wl,--as-needed",
          f"-Wl,-rpath,{OUT}", f"-Wl,-rpath,{B.BUILD}", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath,$ORIGIN/../../loom_amd/_build"])
    os.replace(tmp, FACADE_BIN)
    return [ORACLE_LIB, KERNEL_LIB, FACADE_BIN]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", type=Path, default=Path("/root/reference"))
    ap.add_argument("--oracle-only", action="store_true",
                    help="only the oracle interpreter over the reference's kernels")
    a = ap.parse_args(argv)
    for p in build(a.ref, oracle_only=a.oracle_only):
        print("built", p, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
