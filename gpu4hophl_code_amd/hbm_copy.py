"""STREAM-style device copy through libhsdl.so: achievable HBM bandwidth, and
the known-traffic kernel (k_copy) that rocprofv3 byte counters are
calibrated against.

usage: python3 -m gpu4hophl_code_amd.hbm_copy [--bytes N] [--iters K]
"""
import argparse
import ctypes
import json

from . import _check, lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=4 << 30)
    ap.add_argument("--iters", type=int, default=10)
    a = ap.parse_args()
    g = ctypes.c_double()
    _check(lib().hsdl_device_copy_bench(a.bytes, a.iters, ctypes.byref(g)))
    print(json.dumps({"copy_bytes": a.bytes, "iters": a.iters, "gbps_read_plus_write": g.value}))


if __name__ == "__main__":
    main()
