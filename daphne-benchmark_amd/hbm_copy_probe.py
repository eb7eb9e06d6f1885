"""Calibration kernel for HBM traffic counters: one device-to-device copy of a
known size (default 1 GiB read + 1 GiB written) through torch's copy kernel.
Profiling its FETCH_SIZE / WRITE_SIZE with rocprofv3 gives the factor between
those counters and bytes moved, which scales the counters read for the NDT
kernels."""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    a = ap.parse_args()
    n = a.mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {a.mib} MiB")


if __name__ == "__main__":
    main()
