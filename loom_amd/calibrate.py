"""Copy kernels of known size, for scaling HBM byte counters.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m loom_amd.calibrate
    rocprofv3 --pmc WRITE_SIZE -- python3 -m loom_amd.calibrate

Each of the three dispatches reads 1 GiB and writes 1 GiB (an elementwise
device-to-device copy of a 2^28-element float32 tensor), so the counter's
KiB per dispatch divided by 2^20 is the factor that turns that counter into
bytes.  torch is only the launcher here.
"""
from __future__ import annotations

import sys

N = 1 << 28  # float32 elements: 1 GiB


def main() -> int:
    import torch
    a = torch.ones(N, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {N * 4} bytes x3", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
