"""Copy kernel of known size for calibrating HBM byte counters.

Run under ``rocprofv3 --pmc FETCH_SIZE`` (and separately ``WRITE_SIZE``): the
device-to-device copy of ``--mib`` MiB reads and writes exactly that many
bytes, so counter / expected gives the scale factor applied to the counters
of the measured kernels.
"""

import argparse

import torch

if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    args = ap.parse_args()
    n = args.mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB x3", flush=True)
