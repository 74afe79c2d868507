#!/usr/bin/env python3
"""Reference run for hardware byte counters: copies a buffer of known size
with the library's plain copy kernel (mtgp_device_copy) a few times.

Under `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) each copy dispatch should
read (write) exactly --mib MiB; the ratio of that to the counter's reading,
in KiB per dispatch, is the factor that turns the counter into bytes.

    rocprofv3 --pmc WRITE_SIZE -d out -- python3 mtgp_amd/pmc_calibrate.py
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import mtgp_amd as M  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    n = args.mib << 20
    src, dst = M.DeviceBuffer(n), M.DeviceBuffer(n)
    for _ in range(args.reps):
        M.device_copy(dst, src, n)
    M.synchronize()
    print("copied %d bytes (%d KiB) per dispatch, %d dispatches" % (n, n >> 10, args.reps), flush=True)


if __name__ == "__main__":
    main()
