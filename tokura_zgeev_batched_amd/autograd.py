"""Differentiable eigenvalues through the library's adjoint kernel.

``eigvals_autograd(M)`` returns the eigenvalues of a (B, n, n) complex128
device tensor as a differentiable op whose backward is the C-ABI entry
tokura_zgeev_batched_backward_gpu.  Gradient convention = torch's: for a real
loss L, ``M.grad`` = dL/dM with dL = Re sum conj(G) dM, which for
L = Re sum conj(g_i) lambda_i is G = sum_i g_i y_i x_i^H / conj(y_i^H x_i).

Two modes:
* ``mode="store"`` (default): the forward runs the Schur-vector kernel and
  keeps T and U (2 n^2 complex per matrix) for the backward.
* ``mode="recompute"``: the forward runs the eigenvalue-only kernel and keeps
  only A; the backward recomputes the Schur form.  The forward uses the
  eigenvalue variant whose Hessenberg and QR sweeps are the backward's (for
  n <= 32 the fused single-bulge variant with the adjoint kernel's lanes per
  matrix, for n > 32 the two-bulge QR after the 256-thread workgroup
  Hessenberg), so both passes return the eigenvalues in the same order.
"""
from __future__ import annotations

import torch

from . import (Handle, TokuraError, _P, _check, _dev_check, _stream_ptr, default_handle, lib, workspace_size,
               MAX_N)

# eigenvalue variants whose arithmetic is the Schur-form kernels' (ids of the
# same-size table; n <= 32: fused single-shift QR, 5 = 16x16, 8 = 32x32;
# n > 32: 20, the two-bulge QR after the 256-thread workgroup Hessenberg), so
# the forward's order is the recompute backward's
_MATCHED_VARIANT = {16: 5, 32: 8, 64: 20}


def _adj_nb(n: int) -> int:
    return 16 if n <= 16 else (32 if n <= 32 else 64)


def schur_batched(A_colmajor: torch.Tensor, n: int, batch: int, handle: Handle | None = None, stream=None):
    """Forward with Schur vectors: (eig (B, n), schur (B, 2, n, n) column-major
    T then U, flag (B,))."""
    h = handle or default_handle()
    dev = A_colmajor.device
    eig = torch.empty((batch, n), dtype=torch.complex128, device=dev)
    schur = torch.empty((batch, 2, n, n), dtype=torch.complex128, device=dev)
    flag = torch.empty((batch,), dtype=torch.int8, device=dev)
    rc = lib().tokura_zgeev_batched_schur_gpu(h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(eig.data_ptr()),
                                              _P(schur.data_ptr()), _P(flag.data_ptr()), _stream_ptr(stream))
    _check(rc, "tokura_zgeev_batched_schur_gpu")
    return eig, schur, flag


def backward_batched(n: int, batch: int, gbar: torch.Tensor, schur: torch.Tensor | None = None,
                     A_colmajor: torch.Tensor | None = None, flag: torch.Tensor | None = None,
                     handle: Handle | None = None, stream=None):
    """dL/dA (B, n, n) column-major for cotangents gbar (B, n)."""
    h = handle or default_handle()
    dev = gbar.device
    gbar = gbar.to(torch.complex128).contiguous()
    dA = torch.empty((batch, n, n), dtype=torch.complex128, device=dev)
    if flag is None:
        flag = torch.zeros((batch,), dtype=torch.int8, device=dev)
    rc = lib().tokura_zgeev_batched_backward_gpu(
        h.ptr, n, batch, _P(A_colmajor.data_ptr()) if A_colmajor is not None else None,
        _P(schur.data_ptr()) if schur is not None else None, _P(gbar.data_ptr()), _P(dA.data_ptr()),
        _P(flag.data_ptr()), _stream_ptr(stream))
    _check(rc, "tokura_zgeev_batched_backward_gpu")
    return dA, flag


_RECOMPUTE_HANDLE = None


def recompute_handle() -> Handle:
    """Handle whose eigenvalue variants share the arithmetic of the
    recompute backward's Schur form (same Hessenberg, same QR sweeps), so the
    forward's eigenvalue order is the order the backward differentiates."""
    global _RECOMPUTE_HANDLE
    if _RECOMPUTE_HANDLE is None:
        h = Handle()
        for n in range(1, MAX_N + 1):
            h.set_variant(n, _MATCHED_VARIANT[_adj_nb(n)])
        _RECOMPUTE_HANDLE = h
    return _RECOMPUTE_HANDLE


class _Eigvals(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, mode):
        _dev_check(M, "M", torch.complex128)
        B, n, _ = M.shape
        A = M.transpose(-1, -2).contiguous()
        if mode == "store":
            eig, schur, flag = schur_batched(A, n, B)
            ctx.save_for_backward(schur, flag)
        elif mode == "recompute":
            h = recompute_handle()
            eig = torch.empty((B, n), dtype=torch.complex128, device=M.device)
            flag = torch.empty((B,), dtype=torch.int8, device=M.device)
            work = torch.empty((workspace_size(n, B),), dtype=torch.uint8, device=M.device)
            rc = lib().tokura_zgeev_batched_gpu(h.ptr, n, B, _P(A.data_ptr()), _P(eig.data_ptr()),
                                                _P(work.data_ptr()), _P(flag.data_ptr()), _stream_ptr())
            _check(rc, "tokura_zgeev_batched_gpu")
            ctx.save_for_backward(A, flag)
        else:
            raise TokuraError(f"unknown mode {mode!r}")
        ctx.mode = mode
        ctx.n = n
        ctx.B = B
        ctx.mark_non_differentiable(flag)
        return eig, flag

    @staticmethod
    def backward(ctx, geig, gflag):
        saved, flag = ctx.saved_tensors
        if geig is None:
            return None, None
        if ctx.mode == "store":
            dA, _ = backward_batched(ctx.n, ctx.B, geig, schur=saved, flag=flag.clone())
        else:
            dA, _ = backward_batched(ctx.n, ctx.B, geig, A_colmajor=saved, flag=torch.empty_like(flag),
                                     handle=recompute_handle())
        return dA.transpose(-1, -2), None


def eigvals_autograd(M: torch.Tensor, mode: str = "store"):
    """Differentiable eigenvalues of (B, n, n) complex128 device matrices.

    Returns (eig (B, n), flag (B,)); eigenvalue i of matrix b is the i-th
    diagonal entry of its Schur form, the order of tokura_zgeev_batched_gpu."""
    squeeze = M.dim() == 2
    if squeeze:
        M = M.unsqueeze(0)
    eig, flag = _Eigvals.apply(M, mode)
    if squeeze:
        return eig[0], flag[0]
    return eig, flag
