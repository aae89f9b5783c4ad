"""Render a frame of the smallpt9 path tracer on the GPU and write it as a
binary PPM (the visual end-to-end check the reference demo provides with its
progressive renderer, smallpt/vm/vm.cpp:287-400).

    python -m loom_amd.image out.ppm [--w 512 --h 384 --spp 16]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def to_bytes(img: np.ndarray) -> np.ndarray:
    """Radiance to 8-bit sRGB-like values: clamp to [0, 1], gamma 1/2.2,
    round (smallpt's toInt)."""
    x = np.clip(img.astype(np.float64), 0.0, 1.0)
    return (np.power(x, 1.0 / 2.2) * 255.0 + 0.5).astype(np.uint8)


def write_ppm(path, img: np.ndarray) -> None:
    """img: (h, w, 3) radiance with row 0 at the bottom (as rendered)."""
    h, w, _ = img.shape
    data = to_bytes(img[::-1])  # PPM rows go top to bottom
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode())
        f.write(data.tobytes())


def render(w: int, h: int, spp: int, **params) -> np.ndarray:
    import loom_amd
    from loom_amd import workloads
    ins = workloads.pt_inputs(w, h, spp)
    lx, ly, lz = loom_amd.hybrid_function("smallpt9_f32", **params)(*ins)
    return workloads.pt_image(lx, ly, lz, w, h, spp)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--w", type=int, default=512)
    ap.add_argument("--h", type=int, default=384)
    ap.add_argument("--spp", type=int, default=16)
    a = ap.parse_args(argv)
    write_ppm(a.out, render(a.w, a.h, a.spp))
    print(f"wrote {a.out} ({a.w}x{a.h}, {a.spp} spp)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
