"""A plain device copy of known size, for scaling HBM byte counters.

Run under `rocprofv3 --pmc FETCH_SIZE` (then `WRITE_SIZE`): the copy kernel
reads and writes exactly BYTES bytes per launch, so counter / BYTES is the
factor by which the counter over- or under-reads on this GPU.
"""
import sys

import torch

BYTES = 1 << 30  # 1 GiB read + 1 GiB written per launch


def main():
    n = BYTES // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {BYTES} bytes x 3", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
