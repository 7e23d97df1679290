"""A plain device-to-device copy of a known size.

Profiling counters such as rocprofv3's FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) are read on this copy to get the factor between what the counters
report and the bytes actually moved; the same factor then scales the counters
of the solver kernels.

    python -m sovlp.copy_probe [MiB]     (run from sovereign-moe_amd/)
"""

from __future__ import annotations

import sys


def main(mib: int = 256, reps: int = 3) -> None:
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB x {reps}: each copy reads {n} B and writes {n} B", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 256)
