"""Known-traffic device kernels through libhsdl.so, for the rooflines and for
calibrating rocprofv3 byte counters: a STREAM-style copy (k_copy: achievable
HBM bandwidth, 16 B vector accesses) and, with --gather N, N random 4-byte
loads from the --bytes array (k_gather_bench: each load moves a whole
sector).

usage: python3 -m gpu4hophl_code_amd.hbm_copy [--bytes N] [--iters K] [--gather N]
"""
import argparse
import ctypes
import json

from . import _check, lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=4 << 30)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--gather", type=int, default=0)
    a = ap.parse_args()
    g = ctypes.c_double()
    if a.gather:
        _check(lib().hsdl_device_gather_bench(a.bytes, a.gather, ctypes.byref(g)))
        print(json.dumps({"gather_table_bytes": a.bytes, "reads": a.gather,
                          "algorithmic_bytes": 4 * a.gather, "gbps_algorithmic": g.value}))
        return
    _check(lib().hsdl_device_copy_bench(a.bytes, a.iters, ctypes.byref(g)))
    print(json.dumps({"copy_bytes": a.bytes, "iters": a.iters, "gbps_read_plus_write": g.value}))


if __name__ == "__main__":
    main()
