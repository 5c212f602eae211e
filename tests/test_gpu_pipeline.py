"""Host-streaming pipeline (chunked H2D -> solve -> D2H over several
streams, the reference harness's end-to-end scheme) against host LAPACK."""
import numpy as np
import pytest
import torch

import oracle
from tokura_zgeev_batched_amd.pipeline import pinned_colmajor, stream_eigvals

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,B,chunk,ns", [(20, 5000, 1000, 3), (64, 900, 256, 4), (1, 50, 7, 2)])
def test_stream_eigvals(n, B, chunk, ns):
    rng = np.random.default_rng(n + B)
    M = rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))
    A = pinned_colmajor(B, n)
    A.copy_(torch.from_numpy(np.ascontiguousarray(M.transpose(0, 2, 1)).reshape(B, n * n)))
    print(f"[case] pipeline n={n} B={B}", flush=True)
    eig, flag = stream_eigvals(A, n, chunk=chunk, nstreams=ns)
    assert int(flag.abs().sum()) == 0
    ref, _ = oracle.lapack_zgeev(M)
    assert oracle.max_relerr(ref, eig.numpy()) <= 1e-10
