"""Render a frame of the smallpt9 path tracer on the GPU and write it as a
binary PPM (the visual end-to-end check the reference demo provides with its
progressive renderer, smallpt/vm/vm.cpp:287-400).

    python -m loom_amd.image out.ppm [--w 512 --h 384 --spp 16]
    python -m loom_amd.image out.ppm --progressive 4   # rewrite after each pass

The progressive form renders the frame in passes, each launch taking the
next slice of every pixel's samples, and rewrites the image after each pass
with the mean of the samples so far, reporting the pass time and rate, as
the demo refreshes its window and rate display while outputs arrive
(smallpt/vm/vm.cpp:333-370).  The passes run the same sample threads as one
launch of the whole frame, so the last image equals `render` bit for bit.
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


def render_progressive(w: int, h: int, spp: int, passes: int, on_pass=None,
                       **params) -> np.ndarray:
    """Render in `passes` launches of spp / passes samples per pixel each;
    after pass k, `on_pass(k, image, stats)` gets the mean of the samples
    rendered so far and the pass's run statistics.  Returns the final image
    (equal to `render(w, h, spp)`)."""
    import loom_amd
    from loom_amd import workloads
    if passes < 1 or spp % passes:
        raise ValueError("spp must be a positive multiple of passes")
    per = spp // passes
    npix = w * h
    sid = np.arange(npix * spp, dtype=np.uint32).reshape(npix, spp)
    job = workloads.pt_job(w, h, spp)
    f = loom_amd.hybrid_function("smallpt9_f32", **params)
    samples = np.zeros((3, npix, spp), np.float32)
    img = None
    for k in range(passes):
        lo, hi = k * per, (k + 1) * per
        ids = np.ascontiguousarray(sid[:, lo:hi]).ravel()
        cols = f(ids, np.full(len(ids), job, np.uint32))
        for c in range(3):
            samples[c][:, lo:hi] = np.asarray(cols[c]).reshape(npix, per)
        img = workloads.pt_image(*(np.ascontiguousarray(samples[c][:, :hi]).ravel()
                                   for c in range(3)), w, h, hi)
        if on_pass is not None:
            on_pass(k, img, f.last)
    return img


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--w", type=int, default=512)
    ap.add_argument("--h", type=int, default=384)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--progressive", type=int, default=0, metavar="PASSES",
                    help="render in this many passes, rewriting the image after each")
    a = ap.parse_args(argv)
    if a.progressive:
        per = a.spp // a.progressive

        def show(k, img, st):
            write_ppm(a.out, img)
            rate = a.w * a.h * per / (st.ms_total / 1e3)
            print(f"pass {k + 1}/{a.progressive}: {(k + 1) * per} spp, {st.ms_total:.2f} ms, "
                  f"{rate / 1e9:.3f} G samples/s, active lanes {st.active_lane_fraction:.3f}",
                  flush=True)

        render_progressive(a.w, a.h, a.spp, a.progressive, on_pass=show)
    else:
        write_ppm(a.out, render(a.w, a.h, a.spp))
    print(f"wrote {a.out} ({a.w}x{a.h}, {a.spp} spp)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
