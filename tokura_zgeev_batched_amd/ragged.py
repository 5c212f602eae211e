"""Ragged (mixed-size) batches: the Hessenberg pass (a workgroup per matrix of
n > 32, a wavefront per tile of small matrices) and one persistent QR kernel
for the whole batch.

Wraps tokura_get_zgeev_ragged_worksize / _prepare / _gpu / _graph and
tokuraAutotuneRagged (include/tokura_blas.h).  The reference has no ragged
mode (one n per call, README.md:5).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import Handle, TokuraError, _P, _buf, _check, _dev_check, _stream_ptr, default_handle, lib


def pack(mats):
    """Packs a list of square complex arrays (row/column sense) into the
    ragged layout: (flat complex128 array, int64 offsets)."""
    sizes = [int(m.shape[0]) for m in mats]
    offs = np.zeros(len(mats) + 1, np.int64)
    offs[1:] = np.cumsum(np.array(sizes, np.int64) ** 2)
    flat = np.empty(int(offs[-1]), np.complex128)
    for b, m in enumerate(mats):
        flat[offs[b]:offs[b + 1]] = np.asarray(m, np.complex128).T.reshape(-1)   # column-major
    return flat, offs


class RaggedBatch:
    """A prepared ragged schedule bound to one set of offsets.

    ``offsets``: int64 host array (batch+1), offsets[b] is the start of matrix
    b (column-major) in the flat complex buffer, offsets[b+1]-offsets[b] = n_b^2.
    """

    def __init__(self, offsets, handle: Handle | None = None, device="cuda", stream=None):
        self.h = handle or default_handle()
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        self.batch = len(self.offsets) - 1
        if self.batch <= 0:
            raise TokuraError("empty ragged batch")
        d = np.diff(self.offsets)
        self.sizes = np.rint(np.sqrt(d)).astype(np.int64)
        if np.any(self.sizes ** 2 != d) or np.any(self.sizes < 1) or np.any(self.sizes > 64):
            raise TokuraError("offsets must describe n x n matrices with 1 <= n <= 64")
        self.eig_offsets = np.zeros(self.batch + 1, np.int64)
        self.eig_offsets[1:] = np.cumsum(self.sizes)
        ws = int(lib().tokura_get_zgeev_ragged_worksize(self.batch, self.offsets.ctypes.data_as(_P)))
        if ws >= (1 << 63):
            raise TokuraError("tokura_get_zgeev_ragged_worksize rejected the offsets")
        self.work = torch.empty(ws, dtype=torch.uint8, device=device)
        self.prepare(stream)

    def prepare(self, stream=None):
        rc = lib().tokura_zgeev_ragged_prepare(self.h.ptr, self.batch, self.offsets.ctypes.data_as(_P),
                                               _P(self.work.data_ptr()), self.work.numel(), _stream_ptr(stream))
        _check(rc, "tokura_zgeev_ragged_prepare")

    def _bufs(self, A, eig, flag):
        _dev_check(A, "A", torch.complex128)
        if A.numel() < int(self.offsets.max()):
            raise TokuraError("A is smaller than the offsets require")
        if eig is None:
            eig = torch.empty(int(self.eig_offsets[-1]), dtype=torch.complex128, device=A.device)
        _buf(eig, "eig", torch.complex128, int(self.eig_offsets[-1]), A.device)
        if flag is None:
            flag = torch.empty(self.batch, dtype=torch.int8, device=A.device)
        _buf(flag, "flag", torch.int8, self.batch, A.device)
        return eig, flag

    def solve(self, A: torch.Tensor, eig=None, flag=None, graph: bool = False, stream=None):
        """Eigenvalues of every matrix: returns (eig flat (sum n), flag (batch,))."""
        eig, flag = self._bufs(A, eig, flag)
        f = lib().tokura_zgeev_ragged_graph if graph else lib().tokura_zgeev_ragged_gpu
        rc = f(self.h.ptr, self.batch, _P(A.data_ptr()), _P(eig.data_ptr()), _P(self.work.data_ptr()),
               _P(flag.data_ptr()), _stream_ptr(stream))
        _check(rc, "tokura_zgeev_ragged_graph" if graph else "tokura_zgeev_ragged_gpu")
        return eig, flag

    def schur(self, A: torch.Tensor, stream=None):
        """Forward with Schur vectors: (eig flat (sum n), schur flat (2 * len A),
        flag (batch,)); matrix b's T then U start at 2 * offsets[b]."""
        eig, flag = self._bufs(A, None, None)
        ws = int(lib().tokura_get_zgeev_ragged_schur_worksize(self.batch, self.offsets.ctypes.data_as(_P)))
        if ws >= (1 << 63):
            raise TokuraError("tokura_get_zgeev_ragged_schur_worksize rejected the offsets")
        schur = torch.empty(ws // 16, dtype=torch.complex128, device=A.device)
        rc = lib().tokura_zgeev_ragged_schur_gpu(self.h.ptr, self.batch, _P(A.data_ptr()), _P(eig.data_ptr()),
                                                 _P(schur.data_ptr()), _P(self.work.data_ptr()),
                                                 _P(flag.data_ptr()), _stream_ptr(stream))
        _check(rc, "tokura_zgeev_ragged_schur_gpu")
        return eig, schur, flag

    def backward(self, schur: torch.Tensor, gbar: torch.Tensor, flag: torch.Tensor, stream=None):
        """dL/dA (flat, laid out like A) for flat cotangents gbar (sum n) of
        L = Re sum conj(gbar) lambda; flag from schur()."""
        _buf(schur, "schur", torch.complex128, 2 * int(self.offsets.max()), schur.device)
        gbar = gbar.to(torch.complex128).contiguous()
        _buf(gbar, "gbar", torch.complex128, int(self.eig_offsets[-1]), schur.device)
        _buf(flag, "flag", torch.int8, self.batch, schur.device)
        dA = torch.empty(int(self.offsets.max()), dtype=torch.complex128, device=schur.device)
        rc = lib().tokura_zgeev_ragged_backward_gpu(self.h.ptr, self.batch, _P(schur.data_ptr()), _P(gbar.data_ptr()),
                                                    _P(dA.data_ptr()), _P(self.work.data_ptr()), _P(flag.data_ptr()),
                                                    _stream_ptr(stream))
        _check(rc, "tokura_zgeev_ragged_backward_gpu")
        return dA

    def autotune(self, A: torch.Tensor, stream=None):
        eig, flag = self._bufs(A, None, None)
        rc = lib().tokuraAutotuneRagged(self.h.ptr, self.batch, self.offsets.ctypes.data_as(_P), _P(A.data_ptr()),
                                        _P(eig.data_ptr()), _P(self.work.data_ptr()), self.work.numel(),
                                        _P(flag.data_ptr()), _stream_ptr(stream))
        _check(rc, "tokuraAutotuneRagged")
        return self.config()

    def config(self):
        v = [ctypes.c_int() for _ in range(5)]
        _check(lib().tokuraGetRaggedConfig(self.h.ptr, *[ctypes.byref(x) for x in v]), "tokuraGetRaggedConfig")
        return v[0].value, tuple(x.value for x in v[1:])

    def prepared_config(self):
        """(configuration, Hessenberg grid multiplier) of the schedule prepared in this batch's workspace."""
        c, m = ctypes.c_int(), ctypes.c_int()
        _check(lib().tokuraGetRaggedPrepared(self.h.ptr, _P(self.work.data_ptr()), self.batch, ctypes.byref(c),
                                             ctypes.byref(m)), "tokuraGetRaggedPrepared")
        return c.value, m.value

    def hess_grid(self):
        """Small-class Hessenberg grid multiplier of the handle (workgroups per resident slot)."""
        m = ctypes.c_int()
        _check(lib().tokuraGetRaggedHessGrid(self.h.ptr, ctypes.byref(m)), "tokuraGetRaggedHessGrid")
        return m.value

    def split(self, eig_flat: torch.Tensor):
        """List of per-matrix eigenvalue tensors."""
        return [eig_flat[self.eig_offsets[b]:self.eig_offsets[b + 1]] for b in range(self.batch)]
