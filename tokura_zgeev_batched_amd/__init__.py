"""MI355X-native batched complex eigenvalue solver (n <= 64, complex128).

Python access to the C ABI of ``libtokura_zgeev.so`` (declared in
``include/tokura_blas.h``).  The library is the product; this module only
passes torch device pointers and the current HIP stream through ctypes:

* ``Handle``           tokuraCreate / tokuraDestroy and the tuning controls
* ``zgeev_batched``    same-size batch, the reference call
                       (tokura_zgeev_batched_library/src/tokura_zgeev_batched.cu:212-285)
* ``eigvals``          convenience: (B, n, n) tensor -> (B, n) eigenvalues + flags
* ``zgeev_vectors`` / ``eig``  eigenvalues with right and/or left eigenvectors
                       (ZGEEV with JOBVL/JOBVR = 'V', no balancing)
* ``RaggedBatch``      mixed sizes: Hessenberg launches per LDS size, then one
                       persistent QR kernel for the whole batch (+ hipGraph)
* ``eigvals_autograd`` differentiable eigenvalues (``torch.autograd.Function``
                       whose backward is the library's adjoint kernel)

Layout convention: a (B, n, n) tensor ``M`` means matrices ``M[b]`` in the
usual row/column sense; the library wants each matrix column-major, so the
helpers pass ``M.transpose(-1, -2).contiguous()`` unless told the buffer is
already column-major.  There is no CPU fallback: every call fails loudly when
the HIP library is missing.
"""
from __future__ import annotations

import ctypes
import os

import torch

__all__ = ["lib", "Handle", "zgeev_batched", "eigvals", "workspace_size", "MAX_N", "zgeev_vectors", "eig",
           "RaggedBatch", "eigvals_autograd", "schur_batched", "backward_batched", "backward_recompute"]

MAX_N = 64
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libtokura_zgeev.so")
_LIB = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_SZ = ctypes.c_size_t

_SIGNATURES = {
    "tokuraCreate": ([ctypes.POINTER(_P)], _I),
    "tokuraDestroy": ([_P], _I),
    "tokura_get_zgeeveigenvaluesgetworspacesize": ([_I, _I], _SZ),
    "tokura_zgeev_batched_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P], _I),
    "tokura_zgeev_batched_noalloc_gpu": ([_P, _I, _I, _P, _P, _P, _P], _I),
    "tokuraSetMaxSweeps": ([_P, _I], _I),
    "tokuraSetBalancing": ([_P, ctypes.c_char], _I),
    "tokuraGetLaunchConfig": ([_P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "tokuraAutotune": ([_P, _I, _I, _I, ctypes.c_char_p, _P], _I),
    "tokuraLoadTuning": ([_P, ctypes.c_char_p], _I),
    "tokuraSaveTuning": ([_P, ctypes.c_char_p], _I),
    "tokuraSetVariant": ([_P, _I, _I], _I),
    "tokuraSetStageTiming": ([_P, _I], _I),
    "tokuraGetStageTimes": ([_P, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)], _I),
    "tokura_get_zgeev_ragged_worksize": ([_I, _P], _SZ),
    "tokuraSetRaggedConfig": ([_P, _I], _I),
    "tokuraGetRaggedConfig": ([_P] + [ctypes.POINTER(_I)] * 5, _I),
    "tokuraGetRaggedWaves": ([_P, ctypes.POINTER(_I)], _I),
    "tokuraGetRaggedHessGrid": ([_P, ctypes.POINTER(_I)], _I),
    "tokuraGetRaggedPrepared": ([_P, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "tokuraAutotuneRagged": ([_P, _I, _P, _P, _P, _P, _SZ, _P, _P], _I),
    "tokura_zgeev_ragged_prepare": ([_P, _I, _P, _P, _SZ, _P], _I),
    "tokura_zgeev_ragged_gpu": ([_P, _I, _P, _P, _P, _P, _P], _I),
    "tokura_zgeev_ragged_graph": ([_P, _I, _P, _P, _P, _P, _P], _I),
    "tokura_get_zgeev_schur_worksize": ([_I, _I], _SZ),
    "tokura_zgeev_batched_schur_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P], _I),
    "tokura_zgeev_batched_backward_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P, _P], _I),
    "tokura_zgeev_batched_backward_recompute_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "tokura_get_zgeev_vectors_worksize": ([_I, _I], ctypes.c_size_t),
    "tokura_zgeev_batched_vectors_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "tokura_get_zgeev_ragged_schur_worksize": ([_I, _P], _SZ),
    "tokura_zgeev_ragged_schur_gpu": ([_P, _I, _P, _P, _P, _P, _P, _P], _I),
    "tokura_zgeev_ragged_backward_gpu": ([_P, _I, _P, _P, _P, _P, _P, _P], _I),
    "tokura_get_dgeev_worksize": ([_I, _I], _SZ),
    "tokura_dgeev_batched_gpu": ([_P, _I, _I, _P, _P, _P, _P, _P, _P], _I),
}


def lib_path() -> str:
    return _LIBPATH


def lib():
    """Loads libtokura_zgeev.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(_LIBPATH):
        raise RuntimeError(f"{_LIBPATH} is missing: build it with "
                           "`python -m tokura_zgeev_batched_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(_LIBPATH)
    for name, (args, res) in _SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


class TokuraError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != 0:
        raise TokuraError(f"{what} returned {rc}")


def _stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def workspace_size(n: int, batch: int) -> int:
    return int(lib().tokura_get_zgeeveigenvaluesgetworspacesize(n, batch))


class Handle:
    """Owns a tokurablas_t (one per host thread, as in the reference)."""

    def __init__(self):
        h = _P()
        rc = lib().tokuraCreate(ctypes.byref(h))
        self._h = h
        if rc != 0:
            lib().tokuraDestroy(h)
            self._h = None
            raise TokuraError(f"tokuraCreate returned {rc} (no HIP device?)")

    @property
    def ptr(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().tokuraDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_max_sweeps(self, k: int):
        _check(lib().tokuraSetMaxSweeps(self._h, k), "tokuraSetMaxSweeps")

    def set_balancing(self, job: str):
        """LAPACK ZGEBAL job for this handle's eigenvalue paths: 'N' (default),
        'P', 'S' or 'B'."""
        _check(lib().tokuraSetBalancing(self._h, job.encode()[:1]), "tokuraSetBalancing")

    def launch_config(self, n: int):
        v, l, m = _I(), _I(), _I()
        _check(lib().tokuraGetLaunchConfig(self._h, n, ctypes.byref(v), ctypes.byref(l), ctypes.byref(m)),
               "tokuraGetLaunchConfig")
        return v.value, l.value, m.value

    def set_variant(self, n: int, variant: int):
        _check(lib().tokuraSetVariant(self._h, n, variant), "tokuraSetVariant")

    def autotune(self, n_lo: int = 1, n_hi: int = MAX_N, batch: int = 8192, cache_path: str | None = None,
                 stream=None):
        rc = lib().tokuraAutotune(self._h, n_lo, n_hi, batch,
                                  cache_path.encode() if cache_path else None, _stream_ptr(stream))
        _check(rc, "tokuraAutotune")

    def set_stage_timing(self, on: bool):
        _check(lib().tokuraSetStageTiming(self._h, 1 if on else 0), "tokuraSetStageTiming")

    def stage_times(self):
        """(calls, hessenberg_ms_total, qr_ms_total) since the last read."""
        c, a, b = _I(), ctypes.c_float(), ctypes.c_float()
        _check(lib().tokuraGetStageTimes(self._h, ctypes.byref(c), ctypes.byref(a), ctypes.byref(b)),
               "tokuraGetStageTimes")
        return c.value, a.value, b.value

    def load_tuning(self, cache_path: str) -> int:
        return int(lib().tokuraLoadTuning(self._h, cache_path.encode()))

    def save_tuning(self, cache_path: str):
        """Writes the same-size table and the ragged tuning table, keyed by the device."""
        _check(lib().tokuraSaveTuning(self._h, cache_path.encode()), "tokuraSaveTuning")


_DEFAULT_HANDLE = None


def default_handle() -> Handle:
    global _DEFAULT_HANDLE
    if _DEFAULT_HANDLE is None:
        _DEFAULT_HANDLE = Handle()
    return _DEFAULT_HANDLE


def _dev_check(t: torch.Tensor, name: str, dtype=None):
    if not t.is_cuda:
        raise TokuraError(f"{name} must be a device tensor (no CPU fallback)")
    if dtype is not None and t.dtype != dtype:
        raise TokuraError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise TokuraError(f"{name} must be contiguous")


def _buf(t: torch.Tensor, name: str, dtype, count: int, device=None):
    """Checks a tensor that the library gets as a raw pointer: a contiguous
    device tensor of `dtype` (None: any dtype, `count` in bytes), on `device`
    when given, holding at least `count` elements.  A host tensor, another
    device or a short buffer would otherwise be read or written out of
    bounds by the kernels."""
    _dev_check(t, name, dtype)
    if device is not None and t.device != device:
        raise TokuraError(f"{name} is on {t.device}, the batch on {device}")
    have = t.numel() if dtype is not None else t.numel() * t.element_size()
    if have < count:
        raise TokuraError(f"{name} holds {have} {'elements' if dtype is not None else 'bytes'}, "
                          f"the call needs {count}")


def _shape_ok(n: int, batch: int) -> bool:
    # invalid n or batch: the C entry returns its own code, nothing is launched
    return 1 <= n <= MAX_N and batch > 0


def zgeev_batched(A_colmajor: torch.Tensor, n: int, batch: int, eig: torch.Tensor, work: torch.Tensor,
                  flag: torch.Tensor, handle: Handle | None = None, stream=None) -> int:
    """Raw reference call: buffers already laid out as the C ABI wants
    (A: batch*n*n complex128 column-major per matrix; eig: batch*n; work:
    workspace_size(n, batch) bytes; flag: batch int8).  Returns the C code."""
    h = handle or default_handle()
    if _shape_ok(n, batch):
        d = A_colmajor.device
        _buf(A_colmajor, "A", torch.complex128, batch * n * n)
        _buf(eig, "eig", torch.complex128, batch * n, d)
        _buf(work, "work", None, workspace_size(n, batch), d)
        _buf(flag, "flag", torch.int8, batch, d)
    return int(lib().tokura_zgeev_batched_gpu(h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(eig.data_ptr()),
                                               _P(work.data_ptr()), _P(flag.data_ptr()), _stream_ptr(stream)))


def eigvals(M: torch.Tensor, handle: Handle | None = None, colmajor: bool = False, stream=None):
    """Eigenvalues of a (B, n, n) complex128 device tensor.

    Returns (eig (B, n) complex128, flag (B,) int8).  With colmajor=True the
    last two axes are taken to be (column, row) already."""
    _dev_check(M, "M", torch.complex128)
    if M.dim() == 2:
        M = M.unsqueeze(0)
    B, n, n2 = M.shape
    if n != n2 or not (1 <= n <= MAX_N):
        raise TokuraError(f"matrices must be n x n with 1 <= n <= {MAX_N}, got {tuple(M.shape)}")
    A = M if colmajor else M.transpose(-1, -2).contiguous()
    eig = torch.empty((B, n), dtype=torch.complex128, device=M.device)
    flag = torch.empty((B,), dtype=torch.int8, device=M.device)
    work = torch.empty((workspace_size(n, B),), dtype=torch.uint8, device=M.device)
    _check(zgeev_batched(A, n, B, eig, work, flag, handle, stream), "tokura_zgeev_batched_gpu")
    return eig, flag


def dgeev_workspace_size(n: int, batch: int) -> int:
    return int(lib().tokura_get_dgeev_worksize(n, batch))


def dgeev_batched(A_colmajor: torch.Tensor, n: int, batch: int, wr: torch.Tensor, wi: torch.Tensor,
                  work: torch.Tensor, flag: torch.Tensor, handle: Handle | None = None, stream=None) -> int:
    """Raw tokura_dgeev_batched_gpu: A batch*n*n float64 column-major per
    matrix; wr, wi batch*n float64; work dgeev_workspace_size(n, batch) bytes;
    flag batch int8.  Returns the C code."""
    h = handle or default_handle()
    if _shape_ok(n, batch):
        d = A_colmajor.device
        _buf(A_colmajor, "A", torch.float64, batch * n * n)
        _buf(wr, "wr", torch.float64, batch * n, d)
        _buf(wi, "wi", torch.float64, batch * n, d)
        _buf(work, "work", None, dgeev_workspace_size(n, batch), d)
        _buf(flag, "flag", torch.int8, batch, d)
    return int(lib().tokura_dgeev_batched_gpu(h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(wr.data_ptr()),
                                               _P(wi.data_ptr()), _P(work.data_ptr()), _P(flag.data_ptr()),
                                               _stream_ptr(stream)))


def eigvals_real(M: torch.Tensor, handle: Handle | None = None, stream=None):
    """Eigenvalues of a (B, n, n) float64 device tensor (the DGEEV analogue).

    Returns (w (B, n) complex128, flag (B,) int8): complex conjugate pairs
    are consecutive, positive imaginary part first."""
    _dev_check(M, "M", torch.float64)
    if M.dim() == 2:
        M = M.unsqueeze(0)
    B, n, n2 = M.shape
    if n != n2 or not (1 <= n <= MAX_N):
        raise TokuraError(f"matrices must be n x n with 1 <= n <= {MAX_N}, got {tuple(M.shape)}")
    A = M.transpose(-1, -2).contiguous()
    wr = torch.empty((B, n), dtype=torch.float64, device=M.device)
    wi = torch.empty((B, n), dtype=torch.float64, device=M.device)
    flag = torch.empty((B,), dtype=torch.int8, device=M.device)
    work = torch.empty((dgeev_workspace_size(n, B),), dtype=torch.uint8, device=M.device)
    _check(dgeev_batched(A, n, B, wr, wi, work, flag, handle, stream), "tokura_dgeev_batched_gpu")
    return torch.complex(wr, wi), flag


def zgeev_vectors(A_colmajor: torch.Tensor, n: int, batch: int, eig: torch.Tensor, VL: torch.Tensor | None,
                  VR: torch.Tensor | None, flag: torch.Tensor, handle: Handle | None = None, stream=None,
                  work: torch.Tensor | None = None) -> int:
    """Raw tokura_zgeev_batched_vectors_gpu: A column-major per matrix, eig
    batch*n, VL/VR batch*n*n column-major (either may be None), flag batch
    int8, work (optional, used for one-sided requests) at least
    vectors_workspace_size(n, batch) bytes.  Returns the C code."""
    h = handle or default_handle()
    if _shape_ok(n, batch):
        d = A_colmajor.device
        _buf(A_colmajor, "A", torch.complex128, batch * n * n)
        _buf(eig, "eig", torch.complex128, batch * n, d)
        for V, nm in ((VL, "VL"), (VR, "VR")):
            if V is not None:
                _buf(V, nm, torch.complex128, batch * n * n, d)
        if work is not None:
            _buf(work, "work", None, vectors_workspace_size(n, batch), d)
        _buf(flag, "flag", torch.int8, batch, d)
    return int(lib().tokura_zgeev_batched_vectors_gpu(
        h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(eig.data_ptr()),
        _P(VL.data_ptr()) if VL is not None else None, _P(VR.data_ptr()) if VR is not None else None,
        _P(work.data_ptr()) if work is not None else None, _P(flag.data_ptr()), _stream_ptr(stream)))


def vectors_workspace_size(n: int, batch: int) -> int:
    """Bytes of the optional workspace of a one-sided eigenvector call."""
    return int(lib().tokura_get_zgeev_vectors_worksize(n, batch))


def eig(M: torch.Tensor, left: bool = False, right: bool = True, handle: Handle | None = None, stream=None,
        workspace: bool = True):
    """Eigenvalues and eigenvectors of a (B, n, n) complex128 device tensor.

    Returns (w (B, n), VL (B, n, n) or None, VR (B, n, n) or None, flag (B,)),
    with VR[b, :, i] the right eigenvector of w[b, i] (M[b] v = w v) and
    VL[b, :, i] the left one (u^H M[b] = w u^H), normalised as LAPACK zgeev
    (unit 2-norm, largest component real).  At least one of left/right.
    A one-sided call allocates the optional workspace unless workspace=False
    (same results, slower kernels, no extra memory)."""
    _dev_check(M, "M", torch.complex128)
    if M.dim() == 2:
        M = M.unsqueeze(0)
    B, n, n2 = M.shape
    if n != n2 or not (1 <= n <= MAX_N):
        raise TokuraError(f"matrices must be n x n with 1 <= n <= {MAX_N}, got {tuple(M.shape)}")
    if not (left or right):
        raise TokuraError("eig: ask for left and/or right eigenvectors (eigvals() gives eigenvalues only)")
    A = M.transpose(-1, -2).contiguous()
    w = torch.empty((B, n), dtype=torch.complex128, device=M.device)
    flag = torch.empty((B,), dtype=torch.int8, device=M.device)
    VL = torch.empty((B, n, n), dtype=torch.complex128, device=M.device) if left else None
    VR = torch.empty((B, n, n), dtype=torch.complex128, device=M.device) if right else None
    work = None
    if workspace and not (left and right):
        work = torch.empty(vectors_workspace_size(n, B), dtype=torch.uint8, device=M.device)
    _check(zgeev_vectors(A, n, B, w, VL, VR, flag, handle, stream, work), "tokura_zgeev_batched_vectors_gpu")
    # the library writes column-major: stored[b, i, r] = V(r, i)
    return (w, VL.transpose(-1, -2) if left else None, VR.transpose(-1, -2) if right else None, flag)


from .ragged import RaggedBatch  # noqa: E402
from .autograd import eigvals_autograd, schur_batched, backward_batched, backward_recompute  # noqa: E402
