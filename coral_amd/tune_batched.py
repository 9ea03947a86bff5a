"""Tune the batched GEMM tile kinds and write the table the library loads at start.

For every (ceil log2 m, ceil log2 n, ceil log2 k) bucket of the given range and
each precision, coral_autotune_run times every tile kind of the batched kernel on a
homogeneous batch of the bucket's representative shape (2^(b - 1/2) rounded) and
keeps the fastest; the table is saved as text (one line per bucket:
routine op_a op_b log2m log2n log2k tile) to coral_amd/autotune_table.txt, which
libcoralhip.so reads at initialisation (CORAL_AUTOTUNE_TABLE overrides the path,
"none" disables it).

    python coral_amd/tune_batched.py [--lo 8] [--hi 512] [--out path]

Needs the GPU.  The default range covers m, n, k in [8, 512], the 10,000-problem
benchmark workload.
"""
from __future__ import annotations

import argparse
import ctypes
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def bucket(x: int) -> int:
    b = 0
    while (1 << b) < x:
        b += 1
    return b


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--lo", type=int, default=8)
    ap.add_argument("--hi", type=int, default=512)
    ap.add_argument("--out", default=os.path.join(HERE, "autotune_table.txt"))
    ap.add_argument("--routines", default="sgemm_batched,dgemm_batched")
    args = ap.parse_args()
    os.environ["CORAL_AUTOTUNE_TABLE"] = "none"  # tune from the heuristics, not an old table
    from coral_amd._lib import call, load

    lib = load()
    if lib.coral_init() != 0:
        raise SystemExit(lib.coral_last_error().decode())
    call("coral_autotune_clear")
    buckets = range(bucket(args.lo), bucket(args.hi) + 1)
    t0 = time.time()
    chosen = ctypes.c_int32()
    for routine in args.routines.split(","):
        counts = {}
        for bm in buckets:
            for bn in buckets:
                for bk in buckets:
                    m, n, k = (max(1, round(2.0 ** (b - 0.5))) for b in (bm, bn, bk))
                    call("coral_autotune_run", routine.encode(), 0, 0, m, n, k, ctypes.byref(chosen))
                    counts[chosen.value] = counts.get(chosen.value, 0) + 1
        print(f"{routine}: kinds chosen {dict(sorted(counts.items()))} ({time.time() - t0:.1f} s)", flush=True)
    call("coral_autotune_save", args.out.encode())
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
