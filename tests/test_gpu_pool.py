"""GPU parity of the pool variants of the same-size path (the two-bulge QR on
a persistent workgroup per CU whose LDS is an arena of slots; matrices move
to smaller slots as they deflate) and of the chunk variants (37-39: the LDS
cut into chunks that the deflating matrices compact at every deflation).

The arithmetic is that of variant 20 (the 256-thread Hessenberg and the
two-bulge QR in a fixed LDS buffer), which the oracle's restatement of the
two-bulge sweeps pins (tests/test_gpu_dual.py).  So every pool variant must
give variant 20's eigenvalues and flags bit for bit: on batches large enough
that every CU reuses its slots many times, on a batch smaller than the GPU,
with the sweep cap forcing flag 1, and with non-finite inputs (flag 2).
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk

pytestmark = pytest.mark.gpu

POOL_VARIANTS = list(range(29, 40))


def randn_colmajor(B, n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.complex(torch.randn(B, n * n, generator=g, device="cuda", dtype=torch.float64),
                         torch.randn(B, n * n, generator=g, device="cuda", dtype=torch.float64))


def solve(A, n, B, variant, max_sweeps=None):
    h = tk.Handle()
    assert tk.lib().tokuraSetVariant(h.ptr, n, variant) == 0
    if max_sweeps is not None:
        assert tk.lib().tokuraSetMaxSweeps(h.ptr, max_sweeps) == 0
    eig = torch.empty((B, n), dtype=torch.complex128, device="cuda")
    flag = torch.empty(B, dtype=torch.int8, device="cuda")
    work = torch.empty(tk.workspace_size(n, B), dtype=torch.uint8, device="cuda")
    rc = tk.zgeev_batched(A, n, B, eig, work, flag, h)
    torch.cuda.synchronize()
    h.close()
    assert rc == 0
    return eig, flag


@pytest.mark.parametrize("n", [33, 40, 47, 48, 56, 61, 64])
def test_pool_bitwise_many_rounds(n):
    B = 6000
    A = randn_colmajor(B, n, 3000 + n)
    e20, f20 = solve(A, n, B, 20)
    assert int((f20 != 0).sum()) == 0
    for v in POOL_VARIANTS:
        print(f"[case] pool variant {v} n={n} batch={B}", flush=True)
        e, f = solve(A, n, B, v)
        assert torch.equal(f, f20), v
        assert torch.equal(e, e20), v


@pytest.mark.parametrize("n", [2, 9, 17, 32, 64])
def test_pool_small_batch_and_small_n(n):
    rng = np.random.default_rng(3100 + n)
    B = 7
    M = rng.standard_normal((B, n, n)) + 1j * rng.standard_normal((B, n, n))
    A = torch.tensor(M, device="cuda").transpose(-1, -2).contiguous().reshape(B, n * n)
    ref, _ = oracle.lapack_zgeev(M)
    e20, f20 = solve(A, n, B, 20)
    for v in POOL_VARIANTS:
        print(f"[case] pool variant {v} n={n} batch={B}", flush=True)
        e, f = solve(A, n, B, v)
        assert int((f != 0).sum()) == 0, v
        assert torch.equal(e, e20), v
        assert oracle.max_relerr(ref, e.cpu().numpy()) <= 1e-10, v


@pytest.mark.parametrize("cap", [0, 1, 3])
def test_pool_sweep_cap_flags(cap):
    n, B = 64, 3000
    A = randn_colmajor(B, n, 3200 + cap)
    e20, f20 = solve(A, n, B, 20, max_sweeps=cap)
    assert int((f20 == 1).sum()) > 0
    for v in (29, 31, 34):
        print(f"[case] pool variant {v} sweep cap {cap}", flush=True)
        e, f = solve(A, n, B, v, max_sweeps=cap)
        assert torch.equal(f, f20), v
        assert torch.equal(torch.view_as_real(e).nan_to_num(7.0), torch.view_as_real(e20).nan_to_num(7.0)), v


def test_pool_nonfinite_inputs():
    n, B = 50, 2000
    A = randn_colmajor(B, n, 3300)
    A[5, 17] = float("nan")
    A[1999, 0] = complex(float("inf"), 0.0)
    e20, f20 = solve(A, n, B, 20)
    for v in POOL_VARIANTS:
        print(f"[case] pool variant {v} non-finite", flush=True)
        e, f = solve(A, n, B, v)
        fl = f.cpu().numpy()
        assert fl[5] == 2 and fl[1999] == 2 and int((f != 0).sum()) == 2, v
        assert torch.isnan(e[5]).all() and torch.isnan(e[1999]).all()
        ok = torch.ones(B, dtype=torch.bool, device="cuda")
        ok[5] = ok[1999] = False
        assert torch.equal(e[ok], e20[ok]), v
