"""Workload for profiling runs (rocprofv3) of the LULESH step.

Runs the benchmark workload (Sedov, LULESH 2.0 ICs, 11 regions) for a few
cycles, then, with --copy-bytes, one device-to-device copy of a known size:
the copy is the calibration of the HBM byte counters (FETCH_SIZE,
WRITE_SIZE), which must read and write exactly that many bytes.

    rocprofv3 --kernel-trace --stats -- python3 profile_run.py --nx 150
    rocprofv3 --pmc FETCH_SIZE -- python3 profile_run.py --copy-bytes 1073741824
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import lulesh_amd  # noqa: E402


def device_copy(nbytes: int) -> None:
    """One hipMemcpy device-to-device of nbytes (the calibration kernel)."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) != 0 or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    # hipMemcpyDeviceToDevice = 3
    if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=150)
    ap.add_argument("--cycles", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--mixed", action="store_true",
                    help="also run the mixed-precision step on the same workload")
    args = ap.parse_args()
    if args.cycles > 0:
        for mixed in ((False, True) if args.mixed else (False,)):
            d = lulesh_amd.Domain(args.nx, ic_mode=2, max_cycles=args.warmup + args.cycles,
                                  mixed=mixed)
            d.step(args.warmup)
            rc, n = d.step(args.cycles)
            print(f"lulesh {args.nx}^3{' mixed' if mixed else ''}: {n} cycles, rc {rc}",
                  flush=True)
            d.close()
    if args.copy_bytes > 0:
        device_copy(args.copy_bytes)
        print(f"copied {args.copy_bytes} bytes device to device", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
