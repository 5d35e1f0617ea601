"""Calibration run for hardware memory-traffic counters: a plain streaming
copy kernel of a known size, run a few times.  Under
`rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) each dispatch of the copy
kernel should read (write) exactly --mib MiB; the ratio of that to what the
counter reports is the scale applied to the counter for other kernels.

    python -m namics_amd.hbm_copy [--mib 1024] [--reps 3]
"""
import argparse
import json

import numpy as np

from namics_amd.scf import DeviceBuffer, device_copy


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 8
    src = DeviceBuffer(np.arange(n, dtype=np.float64))
    dst = DeviceBuffer(np.zeros(n))
    for _ in range(a.reps):
        device_copy(dst, src, n * 8)
    print(json.dumps({"copy_bytes_per_dispatch": n * 8, "dispatches": a.reps}), flush=True)
    src.close()
    dst.close()


if __name__ == "__main__":
    main()
