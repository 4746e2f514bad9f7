"""A plain device-to-device copy of known size, used to calibrate the HBM
byte counters (FETCH_SIZE, WRITE_SIZE) of rocprofv3: the counters of this
copy kernel, divided by the bytes it must move, give the scale factor
applied to the transport kernels' counter readings.

usage: python -m gpet_amd.hbm_copy_probe [MiB]
"""
import sys

import torch


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 256
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copied %d MiB x3 (read %d B, write %d B per copy)" % (mib, 4 * n, 4 * n))


if __name__ == "__main__":
    main()
