"""ORACLE — test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The shipped
library (``tokura_zgeev_batched_amd``) never imports, links or calls it.

Contents
--------
* ``method_zgeev``  — the C restatement of the GPU method (zgehd2 + zlahqr,
  eigenvalues only, no balancing), ``zgeev_method.c``.
* ``method_zgeev_dual`` — the same with the two-bulge QR of the GPU's n > 32
  eigenvalue path (two ZLAHQR sweeps per iteration with the shifts of one
  trailing 2x2), ``oracle_zlahqr_dual``.
* ``lapack_zgeev``  — host LAPACK zgeev (balanced, LAPACK default), the
  reference harness's oracle (stream_test_with_MKL/get_eig_MKL.c:26-51).
* ``lapack_hrd_hseqr`` — LAPACK zgehrd + zhseqr('E','N') without balancing.
* ``lapack_rconde`` — LAPACK zgeevx eigenvalue condition numbers (balanc='N').
* ``zgebal``        — ZGEBAL restatement (permutation + power-of-two scaling),
  or host LAPACK zgebal with ``use_lapack=True``.
* ``method_schur`` / ``method_adjoint`` — Schur form and dL/dA from the same
  method, used to check the adjoint kernel.
* ``method_vectors`` — eigenvectors by the same method (ZTREVC on the Schur
  form, back-transform, ZGEEV normalisation); ``lapack_zgeev_vectors`` —
  host LAPACK zgeev('V', 'V') to pin it.
* ``method_dgeev``  — the C restatement of the real path (dgehd2 + dlahqr
  + dlanv2, eigenvalues only, no balancing), ``dgeev_method.c``;
  ``lapack_dgeev`` / ``lapack_dhrd_hseqr`` — host LAPACK dgeev and
  dgehrd + dhseqr; ``dlanv2`` / ``dlahqr`` — the restated steps, or LAPACK's
  own routines with ``use_lapack=True``, to pin them.
* ``match_max_relerr`` — optimal-assignment eigenvalue matching; the
  reference's harness sorts lexicographically instead
  (stream_test_with_MKL/main.cu:16-181), which mis-pairs close eigenvalues.

Conventions: a batch of matrices is a numpy complex128 array ``M`` of shape
(batch, n, n) with ``M[b, r, c]`` = entry (r, c) of matrix b.  The C side
takes column-major matrices, so the helpers transpose the last two axes.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return path


def _lapack_path() -> str:
    import scipy  # noqa: F401  (locates the OpenBLAS that SciPy ships)
    base = os.path.dirname(os.path.dirname(scipy.__file__))
    cands = sorted(glob.glob(os.path.join(base, "scipy.libs", "libscipy_openblas*.so")))
    if not cands:
        raise RuntimeError("SciPy's OpenBLAS (scipy.libs/libscipy_openblas*.so) not found")
    return cands[0]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    I = ctypes.c_int
    L.oracle_zgeev_batched.argtypes = [I, I, P, P, P, P, I]
    L.oracle_zgeev_batched.restype = None
    L.oracle_zgeev_batched_dual.argtypes = [I, I, P, P, P, P, I]
    L.oracle_zgeev_batched_dual.restype = None
    L.oracle_zschur.argtypes = [I, P, P, P, P]
    L.oracle_zschur.restype = I
    L.oracle_zgeev_adjoint.argtypes = [I, P, P, P, P]
    L.oracle_zgeev_adjoint.restype = I
    L.oracle_zgehd2.argtypes = [I, P, I, P, P]
    L.oracle_zgehd2.restype = None
    L.oracle_ztrevc.argtypes = [I, P, P, P]
    L.oracle_ztrevc.restype = None
    L.oracle_zgeev_vectors.argtypes = [I, P, P, P, P]
    L.oracle_zgeev_vectors.restype = I
    L.lapack_zgeev_vectors_batched.argtypes = [I, I, P, P, P, P, P, I]
    L.lapack_zgeev_vectors_batched.restype = I
    L.lapack_open.argtypes = [ctypes.c_char_p]
    L.lapack_open.restype = I
    L.lapack_max_threads.restype = I
    for name in ("lapack_zgeev_batched", "lapack_zhrd_hseqr_batched"):
        f = getattr(L, name)
        f.argtypes = [I, I, P, P, P, I]
        f.restype = I
    L.oracle_zgeev_batched_bal.argtypes = [I, I, P, P, P, I, ctypes.c_char]
    L.oracle_zgeev_batched_bal.restype = None
    L.oracle_zgebal.argtypes = [ctypes.c_char, I, P, I, P, P, P]
    L.oracle_zgebal.restype = I
    L.lapack_zgebal.argtypes = [ctypes.c_char, I, P, P, P, P]
    L.lapack_zgebal.restype = I
    L.lapack_zgeevx_rcond.argtypes = [I, I, P, P, P, P, I]
    L.lapack_zgeevx_rcond.restype = I
    L.oracle_dgeev_batched.argtypes = [I, I, P, P, P, P, I, I]
    L.oracle_dgeev_batched.restype = None
    L.oracle_dgehd2.argtypes = [I, P]
    L.oracle_dgehd2.restype = None
    L.oracle_dlahqr.argtypes = [I, P, I, P, P, I]
    L.oracle_dlahqr.restype = I
    L.oracle_dlanv2.argtypes = [P] * 10
    L.oracle_dlanv2.restype = None
    for name in ("lapack_dgeev_batched", "lapack_dhrd_hseqr_batched"):
        f = getattr(L, name)
        f.argtypes = [I, I, P, P, P, P, I]
        f.restype = I
    L.lapack_dlanv2.argtypes = [P, P]
    L.lapack_dlanv2.restype = I
    L.lapack_dlahqr.argtypes = [I, P, P, P]
    L.lapack_dlahqr.restype = I
    rc = L.lapack_open(_lapack_path().encode())
    if rc != 0:
        raise RuntimeError(f"lapack_open failed ({rc})")
    _LIB = L
    return L


def _colmajor(M: np.ndarray) -> np.ndarray:
    M = np.asarray(M, dtype=np.complex128)
    if M.ndim == 2:
        M = M[None]
    return np.ascontiguousarray(M.transpose(0, 2, 1))


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- real matrices (DGEEV analogue) ----------------------------------------

def _colmajor_real(M: np.ndarray) -> np.ndarray:
    M = np.asarray(M, dtype=np.float64)
    if M.ndim == 2:
        M = M[None]
    return np.ascontiguousarray(M.transpose(0, 2, 1))


def method_dgeev(M, itmax: int = -1, nthreads: int = 0):
    """Eigenvalues of real matrices by the restated real method (DGEHD2 +
    DLAHQR + DLANV2, no balancing): (w = wr + i wi (b, n), info (b,)), info
    0, the DLAHQR INFO, or -1 for a non-finite input."""
    A = _colmajor_real(M)
    b, n, _ = A.shape
    wr = np.zeros((b, n))
    wi = np.zeros((b, n))
    info = np.zeros(b, np.int32)
    lib().oracle_dgeev_batched(n, b, _p(A), _p(wr), _p(wi), _p(info), nthreads, itmax)
    return wr + 1j * wi, info


def lapack_dgeev(M, nthreads: int = 0):
    """Host LAPACK dgeev('N', 'N') (balanced, LAPACK default): (w, info)."""
    A = _colmajor_real(M)
    b, n, _ = A.shape
    wr = np.zeros((b, n))
    wi = np.zeros((b, n))
    info = np.zeros(b, np.int32)
    lib().lapack_dgeev_batched(n, b, _p(A), _p(wr), _p(wi), _p(info), nthreads)
    return wr + 1j * wi, info


def lapack_dhrd_hseqr(M, nthreads: int = 0):
    """Host LAPACK dgehrd + dhseqr('E', 'N'), no balancing: (w, info)."""
    A = _colmajor_real(M)
    b, n, _ = A.shape
    wr = np.zeros((b, n))
    wi = np.zeros((b, n))
    info = np.zeros(b, np.int32)
    lib().lapack_dhrd_hseqr_batched(n, b, _p(A), _p(wr), _p(wi), _p(info), nthreads)
    return wr + 1j * wi, info


def dgehd2(M):
    """DGEHD2 restatement on one real matrix: H (n x n, zeros below the sub-diagonal)."""
    A = _colmajor_real(M)[0].copy()
    lib().oracle_dgehd2(A.shape[0], _p(A))
    return A.T.copy()


def dlahqr(H, use_lapack: bool = False):
    """Eigenvalues of a real upper Hessenberg matrix by the DLAHQR
    restatement or by LAPACK's DLAHQR itself: (wr + i wi, info)."""
    A = _colmajor_real(H)[0].copy()
    n = A.shape[0]
    wr = np.zeros(n)
    wi = np.zeros(n)
    if use_lapack:
        info = lib().lapack_dlahqr(n, _p(A), _p(wr), _p(wi))
    else:
        info = lib().oracle_dlahqr(n, _p(A), n, _p(wr), _p(wi), -1)
    return wr + 1j * wi, info


def dlanv2(a, b, c, d, use_lapack: bool = False):
    """DLANV2 restatement (or LAPACK's): returns
    (a, b, c, d, rt1r, rt1i, rt2r, rt2i, cs, sn) as doubles."""
    if use_lapack:
        abcd = np.array([a, b, c, d], np.float64)
        out = np.zeros(6)
        lib().lapack_dlanv2(_p(abcd), _p(out))
        return tuple(abcd) + tuple(out)
    v = [ctypes.c_double(x) for x in (a, b, c, d)] + [ctypes.c_double() for _ in range(6)]
    lib().oracle_dlanv2(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def method_zgeev_dual(M, itmax: int = -1, with_sweeps: bool = False):
    """Eigenvalues by the restated two-bulge QR of the GPU eigenvalue path
    (oracle_zlahqr_dual after oracle_zgehd2): (eig[b, i], flag[b])."""
    A = _colmajor(M)
    b, n, _ = A.shape
    eig = np.zeros((b, n), np.complex128)
    flag = np.zeros(b, np.int8)
    sweeps = np.zeros(b, np.int32)
    lib().oracle_zgeev_batched_dual(n, b, _p(A), _p(eig), _p(flag), _p(sweeps), itmax)
    if with_sweeps:
        return eig, flag, sweeps
    return eig, flag


def method_zgeev(M, itmax: int = -1, with_sweeps: bool = False, balance: str = "N"):
    """Eigenvalues by the restated GPU method; returns (eig[b, i], flag[b]).
    balance: ZGEBAL job applied first ('N' none, 'P', 'S', 'B')."""
    A = _colmajor(M)
    b, n, _ = A.shape
    eig = np.zeros((b, n), np.complex128)
    flag = np.zeros(b, np.int8)
    if balance != "N":
        lib().oracle_zgeev_batched_bal(n, b, _p(A), _p(eig), _p(flag), itmax, balance.encode())
        return eig, flag
    sweeps = np.zeros(b, np.int32)
    lib().oracle_zgeev_batched(n, b, _p(A), _p(eig), _p(flag), _p(sweeps), itmax)
    if with_sweeps:
        return eig, flag, sweeps
    return eig, flag


def hessenberg(M):
    """zgehd2 restatement on one matrix; returns H (n x n, zeros below sub-diagonal)."""
    A = _colmajor(M)[0].copy()
    n = A.shape[0]
    lib().oracle_zgehd2(n, _p(A), n, None, None)
    return A.T.copy()


def method_schur(M):
    """Full Schur form A = U T U^H by the restated method: (T, U, w, info)."""
    A = _colmajor(M)[0]
    n = A.shape[0]
    T = np.zeros((n, n), np.complex128)
    U = np.zeros((n, n), np.complex128)
    w = np.zeros(n, np.complex128)
    info = lib().oracle_zschur(n, _p(A), _p(T), _p(U), _p(w))
    return T.T.copy(), U.T.copy(), w, info


def method_adjoint(M, g):
    """dL/dA for L = Re sum conj(g_i) lambda_i; g is in Schur-diagonal order.

    Returns (G, w) with w the eigenvalues in that order."""
    A = _colmajor(M)[0]
    n = A.shape[0]
    g = np.ascontiguousarray(np.asarray(g, np.complex128))
    G = np.zeros((n, n), np.complex128)
    w = np.zeros(n, np.complex128)
    lib().oracle_zgeev_adjoint(n, _p(A), _p(g), _p(G), _p(w))
    return G.T.copy(), w


def method_vectors(M, left: bool = True, right: bool = True):
    """Eigenvalues and eigenvectors by the restated GPU method (Schur form,
    ZTREVC, back-transform, ZGEEV normalisation) for a batch.

    Returns (w (b, n), VL (b, n, n) or None, VR (b, n, n) or None, info (b,)),
    with V[b, :, i] the eigenvector of w[b, i] (row/column sense)."""
    A = _colmajor(M)
    b, n, _ = A.shape
    w = np.zeros((b, n), np.complex128)
    VL = np.zeros((b, n, n), np.complex128) if left else None
    VR = np.zeros((b, n, n), np.complex128) if right else None
    info = np.zeros(b, np.int32)
    for j in range(b):
        wj = np.zeros(n, np.complex128)
        vl = np.zeros((n, n), np.complex128) if left else None
        vr = np.zeros((n, n), np.complex128) if right else None
        info[j] = lib().oracle_zgeev_vectors(n, _p(np.ascontiguousarray(A[j])), _p(wj),
                                             _p(vl) if left else None, _p(vr) if right else None)
        w[j] = wj
        if left:
            VL[j] = vl.T     # C wrote column-major: vl[i] is column i
        if right:
            VR[j] = vr.T
    return w, VL, VR, info


def lapack_zgeev_vectors(M, nthreads: int = 0):
    """Host LAPACK zgeev('V', 'V'): (w, VL, VR, info), V[b, :, i] pairs w[b, i]."""
    A = _colmajor(M)
    b, n, _ = A.shape
    w = np.zeros((b, n), np.complex128)
    VL = np.zeros((b, n, n), np.complex128)
    VR = np.zeros((b, n, n), np.complex128)
    info = np.zeros(b, np.int32)
    lib().lapack_zgeev_vectors_batched(n, b, _p(A), _p(w), _p(VL), _p(VR), _p(info), nthreads)
    return w, VL.transpose(0, 2, 1).copy(), VR.transpose(0, 2, 1).copy(), info


def zgebal(M, job: str = "B", use_lapack: bool = False):
    """Balance one matrix (M[r, c] sense): returns (B, ilo, ihi, scale, info)
    with 0-based ilo/ihi; the oracle restatement, or host LAPACK zgebal when
    use_lapack is set."""
    A = _colmajor(M)[0].copy()
    n = A.shape[0]
    ilo, ihi = ctypes.c_int(), ctypes.c_int()
    scale = np.zeros(n, np.float64)
    if use_lapack:
        info = lib().lapack_zgebal(job.encode(), n, _p(A), ctypes.byref(ilo), ctypes.byref(ihi), _p(scale))
        lo, hi = ilo.value - 1, ihi.value - 1
        scale = np.where((np.arange(n) < lo) | (np.arange(n) > hi), scale - 1.0, scale)  # 0-based indices
    else:
        info = lib().oracle_zgebal(job.encode(), n, _p(A), n, ctypes.byref(ilo), ctypes.byref(ihi), _p(scale))
        lo, hi = ilo.value, ihi.value
    return A.T.copy(), lo, hi, scale, info


def lapack_zgeev(M, nthreads: int = 0):
    A = _colmajor(M)
    b, n, _ = A.shape
    eig = np.zeros((b, n), np.complex128)
    info = np.zeros(b, np.int32)
    lib().lapack_zgeev_batched(n, b, _p(A), _p(eig), _p(info), nthreads)
    return eig, info


def lapack_hrd_hseqr(M, nthreads: int = 0):
    A = _colmajor(M)
    b, n, _ = A.shape
    eig = np.zeros((b, n), np.complex128)
    info = np.zeros(b, np.int32)
    lib().lapack_zhrd_hseqr_batched(n, b, _p(A), _p(eig), _p(info), nthreads)
    return eig, info


def lapack_rconde(M, nthreads: int = 0):
    """(eigenvalues, reciprocal condition numbers) of the unbalanced matrices."""
    A = _colmajor(M)
    b, n, _ = A.shape
    w = np.zeros((b, n), np.complex128)
    rc = np.zeros((b, n), np.float64)
    info = np.zeros(b, np.int32)
    lib().lapack_zgeevx_rcond(n, b, _p(A), _p(w), _p(rc), _p(info), nthreads)
    return w, rc, info


def max_threads() -> int:
    return int(lib().lapack_max_threads())


def match(ref: np.ndarray, got: np.ndarray):
    """Optimal assignment of got[i] to ref[i] (per row); returns the
    permutation perm such that got[perm] pairs with ref."""
    from scipy.optimize import linear_sum_assignment
    cost = np.abs(ref[:, None] - got[None, :])
    r, c = linear_sum_assignment(cost)
    perm = np.empty_like(c)
    perm[r] = c
    return perm


def relerr_matched(ref: np.ndarray, got: np.ndarray, scale: np.ndarray | None = None):
    """Per-eigenvalue relative error after optimal matching, for a batch.

    ref, got: (batch, n).  The error of eigenvalue i is |got - ref| / |ref|
    (absolute where ref == 0, as stream_test_with_MKL/main.cu:112-137 does).
    When ``scale`` (batch, n) is given the error is divided by it (used for
    condition scaling).  Returns an array (batch, n)."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    out = np.zeros(ref.shape, np.float64)
    for b in range(ref.shape[0]):
        perm = match(ref[b], got[b])
        g = got[b][perm]
        den = np.abs(ref[b])
        den = np.where(den == 0, 1.0, den)
        e = np.abs(g - ref[b]) / den
        if scale is not None:
            e = e / scale[b]
        out[b] = e
    return out


def max_relerr(ref, got) -> float:
    e = relerr_matched(ref, got)
    return float(e.max()) if e.size else 0.0


def condscaled_err(M, got):
    """Condition-scaled eigenvalue errors of one batch of same-size matrices.

    ref = LAPACK zgeevx eigenvalues of the unbalanced matrices (the GPU
    method does not balance), matched optimally to `got`; error of
    eigenvalue i = |got_i - ref_i| * rconde_i / (eps * ||A||_F), i.e. the
    error in units of the first-order perturbation bound kappa_i eps ||A||.
    Returns an array (batch, n)."""
    M = np.asarray(M, np.complex128)
    if M.ndim == 2:
        M = M[None]
    w, rc, info = lapack_rconde(M)
    eps = np.finfo(np.float64).eps
    out = np.zeros(w.shape)
    for b in range(M.shape[0]):
        perm = match(w[b], np.asarray(got[b]))
        g = np.asarray(got[b])[perm]
        out[b] = np.abs(g - w[b]) * rc[b] / (eps * np.linalg.norm(M[b]))
    return out
