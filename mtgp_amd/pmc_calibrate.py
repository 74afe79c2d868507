#!/usr/bin/env python3
"""Reference runs for hardware counters and the HBM ceiling.

Default: copies a buffer of known size with the library's plain copy kernel
(mtgp_device_copy) a few times.  Under `rocprofv3 --pmc FETCH_SIZE` (or
WRITE_SIZE) each copy dispatch should read (write) exactly --mib MiB; the
ratio of that to the counter's reading, in KiB per dispatch, is the factor
that turns the counter into bytes.

    rocprofv3 --pmc WRITE_SIZE -d out -- python3 mtgp_amd/pmc_calibrate.py

--ceiling: times (HIP events) the plain streaming-store kernel
(mtgp_device_fill, the generation kernels' store pattern without the
generator) and the copy kernel over a 16 GiB buffer and prints GB/s.
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import mtgp_amd as M  # noqa: E402


def ceiling(gib, reps):
    n = gib << 30
    buf = M.DeviceBuffer(n)
    half = n // 2
    s = M.Stream()
    e0, e1 = M.Event(), M.Event()
    for name, fn, moved in (("store (mtgp_device_fill)", lambda: M.device_fill(buf, 0x5A5A5A5A, n, s), n),
                            ("copy (mtgp_device_copy, read+write)",
                             lambda: M.device_copy(buf.address(half), buf, half, s), 2 * half)):
        fn()
        s.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        s.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        print("%-40s %8.1f GB/s  (%d bytes per launch, %.3f ms)" % (name, moved / ms / 1e6, moved, ms), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--ceiling", action="store_true")
    ap.add_argument("--gib", type=int, default=16)
    args = ap.parse_args()
    if args.ceiling:
        return ceiling(args.gib, max(args.reps, 5))
    n = args.mib << 20
    src, dst = M.DeviceBuffer(n), M.DeviceBuffer(n)
    for _ in range(args.reps):
        M.device_copy(dst, src, n)
    M.synchronize()
    print("copied %d bytes (%d KiB) per dispatch, %d dispatches" % (n, n >> 10, args.reps), flush=True)


if __name__ == "__main__":
    main()
