"""Writes tests/golden/zgeev_golden.npz and dgeev_golden.npz: seeded inputs
and host LAPACK outputs (complex zgeev, and real dgeev for the real path).

The reference ships no fixed vectors (its tests draw from std::random_device,
stream_test_with_MKL/main.cu:182-202), so the fixtures are generated here
from host LAPACK (OpenBLAS zgeev, balanced, and zgehrd+zhseqr, unbalanced):
    python tests/golden/make_golden.py
Matrices are stored as M[b, r, c] (row/column sense, not memory order).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

SIZES = [1, 2, 3, 4, 5, 7, 8, 13, 16, 17, 24, 31, 32, 33, 48, 63, 64]
SEED = 20241203


def main():
    rng = np.random.default_rng(SEED)
    out = {}
    for n in SIZES:
        b = 3 if n < 48 else 2
        M = rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))
        e_geev, info1 = oracle.lapack_zgeev(M)
        e_hs, info2 = oracle.lapack_hrd_hseqr(M)
        assert not info1.any() and not info2.any()
        out[f"A_{n}"] = M
        out[f"eig_zgeev_{n}"] = e_geev
        out[f"eig_hseqr_{n}"] = e_hs
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "zgeev_golden.npz")
    np.savez_compressed(path, sizes=np.array(SIZES), seed=np.array(SEED), **out)
    print("wrote", path, os.path.getsize(path), "bytes")
    # real matrices: dgeev (balanced) and dgehrd + dhseqr (unbalanced)
    rng = np.random.default_rng(SEED + 1)
    out = {}
    for n in SIZES:
        b = 3 if n < 48 else 2
        M = rng.standard_normal((b, n, n))
        e_geev, info1 = oracle.lapack_dgeev(M)
        e_hs, info2 = oracle.lapack_dhrd_hseqr(M)
        assert not info1.any() and not info2.any()
        out[f"A_{n}"] = M
        out[f"eig_dgeev_{n}"] = e_geev
        out[f"eig_hseqr_{n}"] = e_hs
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dgeev_golden.npz")
    np.savez_compressed(path, sizes=np.array(SIZES), seed=np.array(SEED + 1), **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
