"""Streams a buffer of known size through the library's plain copy kernel
(gps_selftest_copy) so HBM byte counters (FETCH_SIZE / WRITE_SIZE) can be
calibrated against an exactly known traffic: 1 GiB read + 1 GiB written per
launch, 4 launches.  Usage: python -m gps_amd.copyprobe"""
import sys

import numpy as np

import gps_amd as G


def main() -> int:
    n = 1 << 30
    a = G.DeviceArray((n,), np.uint8, zero=True)
    b = G.DeviceArray((n,), np.uint8)
    for _ in range(4):
        G.check(G.lib.gps_selftest_copy(b.ptr, a.ptr, n, None), "copy")
    G.check(G.lib.gps_stream_synchronize(None), "sync")
    print("copied 4 x 1 GiB", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
