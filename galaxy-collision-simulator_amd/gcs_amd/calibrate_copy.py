"""Plain device-to-device copy of a known size, for scaling rocprofv3's
FETCH_SIZE / WRITE_SIZE counters to bytes:

    rocprofv3 --pmc FETCH_SIZE -f csv -d OUT -o copy -- \
        python3 -m gcs_amd.calibrate_copy --mib 256

The copy kernel reads and writes exactly MiB * 2^20 bytes per launch; the
counter's KiB per dispatch divided by that size is the scale factor.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(a.reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert bool((dst[-8:] == src[-8:]).all())
    print("copied %d bytes x %d" % (4 * n, a.reps), flush=True)


if __name__ == "__main__":
    main()
