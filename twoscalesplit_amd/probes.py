"""Device memory probes from the command line: the streaming-copy rate
(tss_bandwidth_probe, 16-byte reads and writes) and sparse-read rates
(tss_gather_probe, one 8-byte read per stride) -- the reference workloads of
the HBM counter calibration. Prints one JSON line.

    python -m twoscalesplit_amd.probes --copy-mb 256 --gather-mb 2048 --strides 128 64 32
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys

from . import _native as N


def copy_rate(mb: int, reps: int = 3) -> float:
    g = C.c_double()
    N.check(N.lib().tss_bandwidth_probe(mb << 20, reps, C.byref(g)), "bandwidth probe")
    return g.value


def gather_rate(mb: int, stride: int, reps: int = 1) -> float:
    g = C.c_double()
    N.check(N.lib().tss_gather_probe(mb << 20, stride, reps, C.byref(g)), "gather probe")
    return g.value


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--gather-mb", type=int, default=2048)
    ap.add_argument("--strides", type=int, nargs="+", default=[128, 64, 32])
    args = ap.parse_args(argv)
    out = {"copy_mb": args.copy_mb, "copy_gbs": round(copy_rate(args.copy_mb), 1),
           "gather_mb": args.gather_mb,
           "gather_words_gbs": {s: round(gather_rate(args.gather_mb, s), 1)
                                for s in args.strides}}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
