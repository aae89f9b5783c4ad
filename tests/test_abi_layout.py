"""The ctypes mirrors of the C ABI's structures (loom_amd/abi.py) have the
C compiler's layout: every field at the offset, and every structure of the
size, that a C program including include/loom_amd.h sees.  A field added on
one side only would otherwise shift every later field silently."""
from __future__ import annotations

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from loom_amd import abi

ROOT = Path(__file__).resolve().parents[1]
STRUCTS = {
    "loom_kernel_info": abi.KernelInfo,
    "loom_params": abi.Params,
    "loom_run_stats": abi.RunStats,
    "loom_job": abi.Job,
    "loom_server_config": abi.ServerConfig,
    "loom_request": abi.Request,
    "loom_stream_counters": abi.StreamCounters,
    "loom_server_stats": abi.ServerStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "loom_amd.h"',
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, *_ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("layout")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        s, f, v = line.split()
        layout[(s, f)] = int(v)
    return layout


@pytest.mark.parametrize("cname", list(STRUCTS))
def test_struct_layout_matches_c(c_layout, cname):
    py = STRUCTS[cname]
    assert C.sizeof(py) == c_layout[(cname, "sizeof")], cname
    for fname, *_ in py._fields_:
        assert getattr(py, fname).offset == c_layout[(cname, fname)], (cname, fname)
