"""GPU check of tokura_zgeev_batched_noalloc_gpu (the same-size call without
`work`, its scratch in the library's fixed device array): eigenvalues and
flags bit for bit those of tokura_zgeev_batched_gpu on the same matrices, on
two calls in a row (the second smaller, so it runs on what the first left in
the array), for the variants of every size range (fused, split, two-bulge,
pool); the batch limit and the argument checks."""
import numpy as np
import pytest
import torch

import tokura_zgeev_batched_amd as tk

pytestmark = pytest.mark.gpu

P = __import__("ctypes").c_void_p


def packed(n):
    return (n - 1) * (n + 2) // 2 + n


def batch_pair(n, B, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.complex(torch.randn(B, n * n, generator=g, device="cuda", dtype=torch.float64),
                         torch.randn(B, n * n, generator=g, device="cuda", dtype=torch.float64))


def with_work(h, A, n, B):
    eig = torch.empty((B, n), dtype=torch.complex128, device="cuda")
    flag = torch.empty(B, dtype=torch.int8, device="cuda")
    work = torch.empty(tk.workspace_size(n, B), dtype=torch.uint8, device="cuda")
    assert tk.zgeev_batched(A, n, B, eig, work, flag, h) == 0
    return eig, flag


def no_alloc(h, A, n, B):
    eig = torch.full((B, n), complex(7.0, 7.0), dtype=torch.complex128, device="cuda")
    flag = torch.full((B,), 9, dtype=torch.int8, device="cuda")
    rc = tk.lib().tokura_zgeev_batched_noalloc_gpu(h.ptr, n, B, P(A.data_ptr()), P(eig.data_ptr()),
                                                   P(flag.data_ptr()), None)
    assert rc == 0
    return eig, flag


@pytest.mark.parametrize("n,variant", [(4, None), (12, None), (24, None), (40, None), (48, 19), (57, None),
                                       (64, None), (64, 20), (64, 29), (64, 34)])
def test_no_alloc_matches_work_entry_twice(n, variant):
    h = tk.Handle()
    if variant is not None:
        h.set_variant(n, variant)
    for call, B in enumerate((3000, 700)):                 # the second call smaller
        A = batch_pair(n, B, 500 + 7 * n + call)
        A[B // 3, 5 % (n * n)] = float("nan")               # a flag-2 matrix in each
        print(f"[case] no-alloc n={n} variant={variant} call={call} batch={B}", flush=True)
        e1, f1 = with_work(h, A, n, B)
        e2, f2 = no_alloc(h, A, n, B)
        torch.cuda.synchronize()
        assert torch.equal(f1, f2), (n, call)
        ok = f1 != 2
        assert torch.equal(e1[ok], e2[ok]), (n, call)
        assert torch.isnan(e2[~ok]).all()
    h.close()


def test_no_alloc_limit_and_validation():
    h = tk.Handle()
    f = tk.lib().tokura_zgeev_batched_noalloc_gpu
    x = torch.zeros(64, dtype=torch.complex128, device="cuda")
    p = P(x.data_ptr())
    cap = ((128 << 20) - 256) // (16 * packed(64))
    assert f(h.ptr, 64, cap + 1, p, p, p, None) == -9
    assert f(None, 4, 1, p, p, p, None) == -1
    assert f(h.ptr, 0, 1, p, p, p, None) == -2
    assert f(h.ptr, 4, 1, None, p, p, None) == -3
    assert f(h.ptr, 4, 1, p, None, p, None) == -4
    assert f(h.ptr, 4, 0, p, p, p, None) == -6
    assert f(h.ptr, 4, 1, p, p, None, None) == -7
    # the largest batch the array holds
    B = cap
    A = batch_pair(64, B, 77)
    e1, f1 = with_work(h, A, 64, B)
    e2, f2 = no_alloc(h, A, 64, B)
    torch.cuda.synchronize()
    assert torch.equal(f1, f2) and torch.equal(e1, e2)
    h.close()
