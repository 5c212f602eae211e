"""Runtime autotuner front end (replaces the reference's build-time tuner,
tokura_zgeev_batched_library/src/TOKURA_ZHSEQR_BATCHED_GEHRD_TUNING_main.cu:89-242,
which times threads-per-matrix for every n and rebuilds the library).

    python -m tokura_zgeev_batched_amd.tune --out tune_cache.txt [--lo 1 --hi 64 --batch 8192]

times every eligible kernel variant of the same-size path per n on the current
device (tokuraAutotune) and writes the table; point TOKURA_TUNE_CACHE at the
file to have tokuraCreate load it.
"""
from __future__ import annotations

import argparse

from . import Handle


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--lo", type=int, default=1)
    ap.add_argument("--hi", type=int, default=64)
    ap.add_argument("--batch", type=int, default=8192)
    a = ap.parse_args()
    h = Handle()
    h.autotune(a.lo, a.hi, a.batch, a.out)
    for n in range(a.lo, a.hi + 1):
        v, lpm, mpb = h.launch_config(n)
        print(f"n={n:2d} variant={v:2d} lanes/matrix={lpm:2d} matrices/workgroup={mpb:2d}")


if __name__ == "__main__":
    main()
