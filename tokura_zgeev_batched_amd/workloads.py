"""Seeded synthetic workloads of the benchmark configurations (numpy, host).

* ``gaussian(B, n, seed)``            iid complex N(0,1) entries (Re, Im ~ N(0,1))
* ``ragged_sizes(B, lo, hi, seed)``   n ~ U{lo..hi}
* ``companion(coeffs)``               companion matrix of a monic polynomial
* ``hard_spectrum(n, rng, ...)``      A = V diag(lambda) V^-1, lambda in three tight
                                      clusters, kappa(V) set by log-spaced singular values
* ``hard_batch(B, lo, hi, seed)``     the hard-spectra mix: half companions of monic
                                      polynomials with iid complex N(0,1) coefficients,
                                      half clustered similarity transforms

The reference draws U(-1,1) entries from std::random_device
(stream_test_with_MKL/main.cu:182-202); these generators are seeded so every
run and every oracle sees the same matrices.  Matrices are returned as
M[b, r, c] (row/column sense).
"""
from __future__ import annotations

import numpy as np


def gaussian(B: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))


def ragged_sizes(B: int, lo: int, hi: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).integers(lo, hi + 1, B)


def companion(coeffs: np.ndarray) -> np.ndarray:
    """Companion matrix of z^n + c_{n-1} z^{n-1} + ... + c_0 (coeffs = c_0..c_{n-1})."""
    n = len(coeffs)
    C = np.zeros((n, n), np.complex128)
    if n > 1:
        C[np.arange(1, n), np.arange(n - 1)] = 1.0
    C[:, -1] = -np.asarray(coeffs)
    return C


def random_unitary(n: int, rng) -> np.ndarray:
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Q, R = np.linalg.qr(Z)
    d = np.diag(R)
    return Q * (d / np.abs(d))


def hard_spectrum(n: int, rng, cond: float = 1e4, spread: float = 1e-3):
    """A = V diag(lam) V^-1 with lam in 3 clusters of the given spread and
    kappa_2(V) = cond; returns (A, lam)."""
    centers = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    which = np.arange(n) % 3
    lam = centers[which] + spread * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    U, W = random_unitary(n, rng), random_unitary(n, rng)
    sig = np.logspace(0, np.log10(cond), n) if n > 1 else np.ones(1)
    V = (U * sig) @ W.conj().T
    Vinv = (W * (1.0 / sig)) @ U.conj().T
    return V @ np.diag(lam) @ Vinv, lam


def _unitary_batch(k: int, n: int, rng) -> np.ndarray:
    Z = rng.standard_normal((k, n, n)) + 1j * rng.standard_normal((k, n, n))
    Q, R = np.linalg.qr(Z)
    d = np.diagonal(R, axis1=1, axis2=2)
    return Q * (d / np.abs(d))[:, None, :]


def hard_batch(B: int, lo: int, hi: int, seed: int, cond: float = 1e4, spread: float = 1e-3):
    """The hard-spectra mix: B matrices, sizes n ~ U{lo..hi}; even indices are
    companions of monic polynomials with iid complex N(0,1) coefficients, odd
    indices V diag(lam) V^-1 with lam in 3 clusters of the given spread and
    kappa_2(V) = cond.  Returns (mats, kinds) with mats a list of arrays."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, B)
    mats = [None] * B
    kinds = ["companion" if b % 2 == 0 else "clustered" for b in range(B)]
    for n in np.unique(sizes):
        n = int(n)
        idx = np.nonzero(sizes == n)[0]
        comp = idx[idx % 2 == 0]
        clus = idx[idx % 2 == 1]
        if len(comp):
            c = rng.standard_normal((len(comp), n)) + 1j * rng.standard_normal((len(comp), n))
            C = np.zeros((len(comp), n, n), np.complex128)
            if n > 1:
                C[:, np.arange(1, n), np.arange(n - 1)] = 1.0
            C[:, :, -1] = -c
            for j, b in enumerate(comp):
                mats[b] = C[j]
        if len(clus):
            k = len(clus)
            centers = rng.standard_normal((k, 3)) + 1j * rng.standard_normal((k, 3))
            lam = centers[:, np.arange(n) % 3] + spread * (rng.standard_normal((k, n)) +
                                                           1j * rng.standard_normal((k, n)))
            U, W = _unitary_batch(k, n, rng), _unitary_batch(k, n, rng)
            sig = np.logspace(0, np.log10(cond), n) if n > 1 else np.ones(1)
            V = (U * sig[None, None, :]) @ np.conj(W.transpose(0, 2, 1))
            Vinv = (W * (1.0 / sig)[None, None, :]) @ np.conj(U.transpose(0, 2, 1))
            A = (V * lam[:, None, :]) @ Vinv
            for j, b in enumerate(clus):
                mats[b] = A[j]
    return mats, kinds


def _orthogonal_batch(k: int, n: int, rng) -> np.ndarray:
    Q, R = np.linalg.qr(rng.standard_normal((k, n, n)))
    return Q * np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]


def hard_batch_real(B: int, n: int, seed: int, cond: float = 1e4, spread: float = 1e-3):
    """Real hard spectra, B matrices of size n: even indices are companions of
    monic polynomials with iid real N(0,1) coefficients; odd indices
    V blockdiag(...) V^-1 with a real orthogonal-times-log-spaced V
    (kappa_2(V) = cond) and a real-Schur-form core whose spectrum sits in
    three clusters of the given spread: one on the real axis around a real
    centre and one conjugate pair of clusters around c +- i d (2x2 blocks
    [[a, b], [-b, a]] with eigenvalues a +- i b).  Returns (M[B, n, n], kinds)."""
    rng = np.random.default_rng(seed)
    M = np.zeros((B, n, n))
    kinds = ["companion" if b % 2 == 0 else "clustered" for b in range(B)]
    comp = np.arange(0, B, 2)
    clus = np.arange(1, B, 2)
    if n > 1:
        M[comp[:, None], np.arange(1, n)[None, :], np.arange(n - 1)[None, :]] = 1.0
    M[comp, :, -1] = -rng.standard_normal((len(comp), n))
    k = len(clus)
    if k:
        D = np.zeros((k, n, n))
        npair = n // 3                       # conjugate pairs use 2 * npair slots
        nreal = n - 2 * npair
        creal = rng.standard_normal(k)
        cre, cim = rng.standard_normal(k), np.abs(rng.standard_normal(k)) + 0.5
        lam_r = creal[:, None] + spread * rng.standard_normal((k, nreal))
        D[:, np.arange(nreal), np.arange(nreal)] = lam_r
        a = cre[:, None] + spread * rng.standard_normal((k, npair))
        b = cim[:, None] + spread * rng.standard_normal((k, npair))
        r0 = nreal + 2 * np.arange(npair)
        D[:, r0, r0] = a
        D[:, r0 + 1, r0 + 1] = a
        D[:, r0, r0 + 1] = b
        D[:, r0 + 1, r0] = -b
        U, W = _orthogonal_batch(k, n, rng), _orthogonal_batch(k, n, rng)
        sig = np.logspace(0, np.log10(cond), n) if n > 1 else np.ones(1)
        V = (U * sig[None, None, :]) @ W.transpose(0, 2, 1)
        Vinv = (W * (1.0 / sig)[None, None, :]) @ U.transpose(0, 2, 1)
        M[clus] = V @ D @ Vinv
    return M, kinds
