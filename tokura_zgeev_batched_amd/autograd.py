"""Differentiable eigenvalues through the library's adjoint kernel.

``eigvals_autograd(M)`` returns the eigenvalues of a (B, n, n) complex128
device tensor as a differentiable op whose backward is a C-ABI entry
(tokura_zgeev_batched_backward_gpu or _backward_recompute_gpu).  Gradient
convention = torch's: for a real loss L, ``M.grad`` = dL/dM with
dL = Re sum conj(G) dM, which for L = Re sum conj(g_i) lambda_i is
G = sum_i g_i y_i x_i^H / conj(y_i^H x_i).

Two modes:
* ``mode="store"`` (default): the forward runs the Schur-vector kernel and
  keeps T and U (2 n^2 complex per matrix) for the backward.
* ``mode="recompute"``: the forward runs the eigenvalue-only path with the
  handle's own (autotuned) variants and keeps A and the eigenvalues; the
  backward rebuilds the Schur form in a workspace and pairs its eigenvalues
  with the forward's on the device, so any forward variant is paired
  correctly.
"""
from __future__ import annotations

import torch

from . import Handle, TokuraError, _P, _buf, _check, _dev_check, _shape_ok, _stream_ptr, default_handle, lib, \
    workspace_size


def schur_batched(A_colmajor: torch.Tensor, n: int, batch: int, handle: Handle | None = None, stream=None):
    """Forward with Schur vectors: (eig (B, n), schur (B, 2, n, n) column-major
    T then U, flag (B,))."""
    h = handle or default_handle()
    dev = A_colmajor.device
    if _shape_ok(n, batch):
        _buf(A_colmajor, "A", torch.complex128, batch * n * n)
    eig = torch.empty((batch, n), dtype=torch.complex128, device=dev)
    schur = torch.empty((batch, 2, n, n), dtype=torch.complex128, device=dev)
    flag = torch.empty((batch,), dtype=torch.int8, device=dev)
    rc = lib().tokura_zgeev_batched_schur_gpu(h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(eig.data_ptr()),
                                              _P(schur.data_ptr()), _P(flag.data_ptr()), _stream_ptr(stream))
    _check(rc, "tokura_zgeev_batched_schur_gpu")
    return eig, schur, flag


def backward_batched(n: int, batch: int, gbar: torch.Tensor, schur: torch.Tensor, flag: torch.Tensor,
                     handle: Handle | None = None, stream=None):
    """dL/dA (B, n, n) column-major from a stored Schur form (B, 2, n, n)
    and cotangents gbar (B, n); `flag` is the forward's."""
    h = handle or default_handle()
    dev = gbar.device
    gbar = gbar.to(torch.complex128).contiguous()
    if _shape_ok(n, batch):
        _buf(gbar, "gbar", torch.complex128, batch * n)
        _buf(schur, "schur", torch.complex128, batch * 2 * n * n, dev)
        _buf(flag, "flag", torch.int8, batch, dev)
    dA = torch.empty((batch, n, n), dtype=torch.complex128, device=dev)
    rc = lib().tokura_zgeev_batched_backward_gpu(h.ptr, n, batch, None, _P(schur.data_ptr()), _P(gbar.data_ptr()),
                                                 _P(dA.data_ptr()), _P(flag.data_ptr()), _stream_ptr(stream))
    _check(rc, "tokura_zgeev_batched_backward_gpu")
    return dA, flag


def backward_recompute(n: int, batch: int, A_colmajor: torch.Tensor, eig: torch.Tensor, gbar: torch.Tensor,
                       handle: Handle | None = None, stream=None):
    """Recompute-mode dL/dA (B, n, n) column-major: the Schur form is rebuilt
    from A in a workspace, its eigenvalues paired with `eig` (the forward's,
    (B, n), any variant's order) and gbar (B, n) applied in that order.
    Returns (dA, flag) with flag 3 where `eig` is not the spectrum of A."""
    h = handle or default_handle()
    dev = gbar.device
    gbar = gbar.to(torch.complex128).contiguous()
    eig = eig.contiguous()
    if _shape_ok(n, batch):
        _buf(gbar, "gbar", torch.complex128, batch * n)
        _buf(A_colmajor, "A", torch.complex128, batch * n * n, dev)
        _buf(eig, "eig", torch.complex128, batch * n, dev)
    dA = torch.empty((batch, n, n), dtype=torch.complex128, device=dev)
    flag = torch.empty((batch,), dtype=torch.int8, device=dev)
    work = torch.empty((workspace_size(n, batch),), dtype=torch.uint8, device=dev)
    rc = lib().tokura_zgeev_batched_backward_recompute_gpu(
        h.ptr, n, batch, _P(A_colmajor.data_ptr()), _P(eig.data_ptr()), _P(gbar.data_ptr()), _P(dA.data_ptr()),
        _P(work.data_ptr()), _P(flag.data_ptr()), _stream_ptr(stream))
    _check(rc, "tokura_zgeev_batched_backward_recompute_gpu")
    return dA, flag


class _Eigvals(torch.autograd.Function):
    @staticmethod
    def forward(ctx, M, mode):
        _dev_check(M, "M", torch.complex128)
        B, n, _ = M.shape
        A = M.transpose(-1, -2).contiguous()
        if mode == "store":
            eig, schur, flag = schur_batched(A, n, B)
            ctx.save_for_backward(schur, None, flag)
        elif mode == "recompute":
            eig = torch.empty((B, n), dtype=torch.complex128, device=M.device)
            flag = torch.empty((B,), dtype=torch.int8, device=M.device)
            work = torch.empty((workspace_size(n, B),), dtype=torch.uint8, device=M.device)
            rc = lib().tokura_zgeev_batched_gpu(default_handle().ptr, n, B, _P(A.data_ptr()), _P(eig.data_ptr()),
                                                _P(work.data_ptr()), _P(flag.data_ptr()), _stream_ptr())
            _check(rc, "tokura_zgeev_batched_gpu")
            ctx.save_for_backward(A, eig, flag)
        else:
            raise TokuraError(f"unknown mode {mode!r}")
        ctx.mode = mode
        ctx.n = n
        ctx.B = B
        ctx.mark_non_differentiable(flag)
        return eig, flag

    @staticmethod
    def backward(ctx, geig, gflag):
        saved, eig, flag = ctx.saved_tensors
        if geig is None:
            return None, None
        if ctx.mode == "store":
            dA, _ = backward_batched(ctx.n, ctx.B, geig, saved, flag.clone())
        else:
            dA, _ = backward_recompute(ctx.n, ctx.B, saved, eig, geig)
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
