"""Seeded synthetic inputs for the benchmark configurations.

All randomness is splitmix64 (Steele, Lea & Flood 2014) over a counter, so
inputs are identical on every machine and in every language that restates
it; no std distributions (implementation-defined) are used.  The default
seed is 0x4C4F4F4D ("LOOM").
"""
from __future__ import annotations

import numpy as np

SEED = 0x4C4F4F4D
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of splitmix64 started at state `seed`, skipping `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform_int(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    """Integers uniform in [lo, hi] (inclusive) by 32-bit multiply-shift of
    the high word (Lemire's reduction without rejection)."""
    span = hi - lo + 1
    assert 0 < span <= 1 << 32
    top = splitmix64(seed, n) >> np.uint64(32)
    return (np.uint64(lo) + ((top * np.uint64(span)) >> np.uint64(32))).astype(np.uint64)


def fib_inputs(n: int = 1 << 20, seed: int = SEED) -> np.ndarray:
    """fib config: n ∈ [0, 90] (F(90) < 2^64)."""
    return uniform_int(seed, n, 0, 90)


def collatz_inputs(n: int = 1 << 24, seed: int = SEED + 1) -> np.ndarray:
    """Collatz config: start values uniform in [1, 2^32)."""
    return uniform_int(seed, n, 1, (1 << 32) - 1)


def pt_job(w: int, h: int, spp: int) -> int:
    """The `job` input word of the smallpt9 kernel: w<<20 | h<<8 | spp."""
    assert 0 < w < 4096 and 0 < h < 4096 and 0 < spp < 256
    return (w << 20) | (h << 8) | spp


def pt_inputs(w: int = 1024, h: int = 768, spp: int = 64):
    """Path-tracing config: one thread per sample, sample index sid and the
    frame word (identical for every thread of the frame)."""
    n = w * h * spp
    return [np.arange(n, dtype=np.uint32), np.full(n, pt_job(w, h, spp), dtype=np.uint32)]


def pt_image(lx, ly, lz, w: int, h: int, spp: int) -> np.ndarray:
    """Pixel radiance: the mean of each pixel's spp samples, summed in sample
    order in float32 (rows bottom to top as generated)."""
    out = np.zeros((h * w, 3), np.float32)
    for c, col in enumerate((lx, ly, lz)):
        acc = np.zeros(h * w, np.float32)
        s = np.asarray(col, np.float32).reshape(h * w, spp)
        for k in range(spp):
            acc = (acc + s[:, k]).astype(np.float32)
        out[:, c] = acc / np.float32(spp)
    return out.reshape(h, w, 3)


def batch_jobs(njobs: int = 64, seed: int = SEED + 2, scale: int = 1):
    """The mixed batch: `njobs` job specs drawn from `seed`.

    Even jobs are path-trace frames with width and height uniform in
    [64, 1024] and spp uniform in [1, 64]; odd jobs alternate fib and
    Collatz with 2^k threads, k uniform in [12, 22].  `scale` > 1 divides
    every size (for tests).  Returns a list of (kernel name, spec dict)."""
    r = uniform_int(seed, 4 * njobs, 0, (1 << 32) - 1).astype(np.uint64)
    jobs = []
    for j in range(njobs):
        a, b, c, d = (int(x) for x in r[4 * j:4 * j + 4])
        if j % 2 == 0:
            w = max(1, (64 + a % 961) // scale)
            h = max(1, (64 + b % 961) // scale)
            spp = max(1, (1 + c % 64) // scale)
            jobs.append(("smallpt9_f32", {"w": w, "h": h, "spp": spp, "n": w * h * spp}))
        else:
            n = max(1, (1 << (12 + a % 11)) // scale)
            name = "fib_u64" if j % 4 == 1 else "collatz_u64"
            jobs.append((name, {"n": n, "seed": seed + 1000 + j}))
    return jobs


def job_inputs(name: str, spec: dict):
    """Host input columns of one batch job spec."""
    if name.startswith("smallpt9"):
        return pt_inputs(spec["w"], spec["h"], spec["spp"])
    if name.startswith("fib"):
        return [fib_inputs(spec["n"], seed=spec["seed"])]
    return [collatz_inputs(spec["n"], seed=spec["seed"])]


def uniform_real(seed: int, n: int, lo: float, hi: float, dtype=np.float64) -> np.ndarray:
    """Reals uniform in [lo, hi) from the top 53 bits of splitmix64."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return (lo + (hi - lo) * u).astype(dtype)


def cubic_inputs(n: int, dtype=np.float32, seed: int = SEED + 6):
    """Cubic coefficients as the reference's root-finding test draws them
    (tests/vm_root_finding.cpp:73-78): a = ±U[1, B], b, c, d = U[-B, B] with
    B = 100 for float32 and 1e5 for float64."""
    bound = 100.0 if np.dtype(dtype) == np.float32 else 1e5
    a = uniform_real(seed, n, 1.0, bound, dtype)
    sign = np.where(splitmix64(seed + 1, n) & np.uint64(1), -1, 1).astype(dtype)
    return [a * sign] + [uniform_real(seed + 2 + k, n, -bound, bound, dtype) for k in range(3)]


def threefry_inputs(n: int, seed: int = SEED + 7):
    """Counter words (c0, c1) for the Threefry kernel."""
    r = splitmix64(seed, n)
    return [(r & np.uint64(0xFFFFFFFF)).astype(np.uint32), (r >> np.uint64(32)).astype(np.uint32)]
