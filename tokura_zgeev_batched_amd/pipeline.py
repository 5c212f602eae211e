"""Host-streaming pipeline: eigenvalues of a host-resident batch.

The reference's harness measures end to end: 8 streams, chunks of 8192
matrices, each chunk copied host->device, solved and copied back
(stream_test_with_MKL/main.cu:245-296, stream_test/main.cu:76-127).
``stream_eigvals`` does the same on HIP streams: the batch is split into
chunks, chunk c goes to stream c % nstreams, and each stream runs
H2D copy -> tokura_zgeev_batched_gpu -> D2H copy, so copies of one chunk
overlap the solve of another.  Host buffers are pinned (allocated pinned, not
registered after the fact).
"""
from __future__ import annotations

import torch

from . import Handle, TokuraError, _check, default_handle, workspace_size, zgeev_batched


def pinned_colmajor(B: int, n: int) -> torch.Tensor:
    """A pinned host buffer for B column-major n x n complex128 matrices."""
    return torch.empty((B, n * n), dtype=torch.complex128, pin_memory=True)


def stream_eigvals(A_host: torch.Tensor, n: int, chunk: int = 8192, nstreams: int = 4,
                   handle: Handle | None = None):
    """A_host: pinned (B, n*n) complex128, each row one column-major matrix.

    Returns (eig (B, n) pinned host tensor, flag (B,) pinned int8), valid once
    the call returns (it synchronises its streams)."""
    if A_host.is_cuda or not A_host.is_pinned():
        raise TokuraError("A_host must be a pinned host tensor")
    B = A_host.shape[0]
    h = handle or default_handle()
    dev = torch.device("cuda")
    eig = torch.empty((B, n), dtype=torch.complex128, pin_memory=True)
    flag = torch.empty((B,), dtype=torch.int8, pin_memory=True)
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    cs = min(chunk, B)
    bufs = [(torch.empty((cs, n * n), dtype=torch.complex128, device=dev),
             torch.empty((cs, n), dtype=torch.complex128, device=dev),
             torch.empty((cs,), dtype=torch.int8, device=dev),
             torch.empty((workspace_size(n, cs),), dtype=torch.uint8, device=dev)) for _ in range(nstreams)]
    cur = torch.cuda.current_stream()
    for s in streams:
        s.wait_stream(cur)   # device buffers were allocated on the current stream
    for c, start in enumerate(range(0, B, cs)):
        s = streams[c % nstreams]
        dA, dE, dF, dW = bufs[c % nstreams]
        m = min(cs, B - start)
        with torch.cuda.stream(s):
            dA[:m].copy_(A_host[start:start + m], non_blocking=True)
            _check(zgeev_batched(dA, n, m, dE, dW, dF, h, s), "tokura_zgeev_batched_gpu")
            eig[start:start + m].copy_(dE[:m], non_blocking=True)
            flag[start:start + m].copy_(dF[:m], non_blocking=True)
    for s in streams:
        s.synchronize()
    return eig, flag
