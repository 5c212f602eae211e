"""GPU checks of optional balancing (tokuraSetBalancing, the ZGEBAL analogue).

* Graded matrices A = D M D^-1 with D spanning 2^40: with job 'B' (or 'S') every
  kernel family agrees with the eigenvalues of M and with host LAPACK zgeev to
  <= 1e-10; without balancing the error is orders of magnitude larger.
* Permuted block-triangular matrices and zero rows/columns (isolated
  eigenvalues, job 'P' and 'B') against LAPACK zgeev.
* The ragged path with balancing; non-finite input still flags 2; job 'N'
  gives the same bits as a handle that never set balancing.
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rand(rng, b, n):
    return rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))


def graded(rng, b, n, span=40.0):
    M = rand(rng, b, n)
    d = 2.0 ** np.linspace(0, span, n)
    return (d[None, :, None] * M) / d[None, None, :], M


def eigs(M, h):
    e, f = tk.eigvals(torch.tensor(M, device="cuda"), h)
    torch.cuda.synchronize()
    return e.cpu().numpy(), f.cpu().numpy()


# (n, variant): fused, split single-bulge, split two-bulge, and the default table
CASES = [(5, 1), (12, 4), (16, 5), (24, 8), (33, 15), (40, 16), (48, 19), (57, 20), (64, 18), (64, None)]


@pytest.mark.parametrize("n,variant", CASES)
def test_graded_balanced(n, variant):
    rng = np.random.default_rng(5000 + n)
    G, M = graded(rng, 40, n)
    exact = np.stack([np.linalg.eigvals(m) for m in M])
    h = tk.Handle()
    if variant is not None:
        assert tk.lib().tokuraSetVariant(h.ptr, n, variant) == 0
    print(f"[case] balanced graded n={n} variant={variant}", flush=True)
    h.set_balancing("B")
    e, f = eigs(G, h)
    assert not f.any()
    assert oracle.max_relerr(exact, e) <= TOL
    ref, _ = oracle.lapack_zgeev(G)
    assert oracle.max_relerr(ref, e) <= TOL
    h.set_balancing("S")
    e_s, f = eigs(G, h)
    assert oracle.max_relerr(exact, e_s) <= TOL
    h.set_balancing("N")
    e0, f0 = eigs(G, h)
    h.close()
    if n >= 16:
        assert oracle.max_relerr(exact, e0) > 1e-6  # unbalanced: the grading destroys accuracy


def test_isolated_eigenvalues():
    rng = np.random.default_rng(51)
    mats = []
    for _ in range(20):
        n = 20
        T = np.triu(rand(rng, 1, n)[0])
        T[5:14, 5:14] = rand(rng, 1, 9)[0]
        P = rng.permutation(n)
        mats.append(T[np.ix_(P, P)])
    M = np.stack(mats)
    ref, _ = oracle.lapack_zgeev(M)
    for job in ("P", "B"):
        h = tk.Handle()
        h.set_balancing(job)
        print(f"[case] isolated job={job}", flush=True)
        e, f = eigs(M, h)
        h.close()
        assert not f.any()
        assert oracle.max_relerr(ref, e) <= TOL
    Z = rand(rng, 8, 10)
    Z[:, 6, :] = 0
    Z[:, 6, 6] = 2
    Z[:, :, 1] = 0
    Z[:, 1, 1] = -1
    h = tk.Handle()
    h.set_balancing("B")
    e, f = eigs(Z, h)
    h.close()
    ref, _ = oracle.lapack_zgeev(Z)
    assert not f.any() and oracle.max_relerr(ref, e) <= TOL


def test_balancing_off_is_unchanged_and_flags():
    rng = np.random.default_rng(52)
    M = rand(rng, 30, 40)
    h1 = tk.Handle()
    h2 = tk.Handle()
    h2.set_balancing("B")
    h2.set_balancing("N")
    e1, _ = eigs(M, h1)
    e2, _ = eigs(M, h2)
    np.testing.assert_array_equal(e1, e2)
    M[3, 2, 7] = np.nan
    h2.set_balancing("B")
    e, f = eigs(M, h2)
    h1.close()
    h2.close()
    assert f[3] == 2 and np.isnan(e[3]).all()
    assert (np.delete(f, 3) == 0).all()


def test_ragged_balanced():
    rng = np.random.default_rng(53)
    sizes = rng.integers(4, 65, 300)
    mats, exact = [], []
    for n in sizes:
        G, M = graded(rng, 1, int(n))
        mats.append(G[0])
        exact.append(np.linalg.eigvals(M[0]))
    flat, offs = pack(mats)
    h = tk.Handle()
    h.set_balancing("B")
    rb = RaggedBatch(offs, handle=h)
    print("[case] ragged balanced 300", flush=True)
    eig, flag = rb.solve(torch.tensor(flat, device="cuda"), graph=True)
    torch.cuda.synchronize()
    assert int((flag != 0).sum()) == 0
    parts = rb.split(eig)
    worst = max(oracle.max_relerr(exact[b][None], parts[b].cpu().numpy()[None]) for b in range(len(mats)))
    h.close()
    assert worst <= TOL
