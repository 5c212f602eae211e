"""GPU parity of the real path (tokura_dgeev_batched_gpu, the DGEEV analogue:
DGEHD2 + DLAHQR double-shift QR + DLANV2, eigenvalues only, no balancing).

Bars (as the complex path): max relative error <= 1e-10 after optimal
matching against the oracle's restatement and against host LAPACK dgeev;
complex eigenvalues in exact conjugate pairs (positive imaginary part first,
equal real parts); closed forms (triangular input, rotations, roots of
unity, symmetric input); flags for non-finite input and the sweep cap; at
the full 100,000 x 64 size the trace invariant sum(wr) = trace(A) and, on a
sample, the condition-scaled error (<= 100 units of kappa_i eps ||A||_F) and
the error relative to ||A||_F (<= 1e-13) against LAPACK dgeev; hard spectra
(real companions, clustered ill-conditioned real spectra) with the
condition-scaled error <= 100 and no flag.
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.workloads import hard_batch_real

pytestmark = pytest.mark.gpu

TOL = 1e-10


def eigs(M, h=None):
    print(f"[case] real n={M.shape[-1]} batch={M.shape[0]}", flush=True)
    w, f = tk.eigvals_real(torch.tensor(M, device="cuda"), h)
    torch.cuda.synchronize()
    return w.cpu().numpy(), f.cpu().numpy()


def check_pairs(w):
    for row in w:
        j = 0
        while j < len(row):
            if row[j].imag != 0.0:
                assert row[j].imag > 0 and row[j + 1].imag == -row[j].imag and row[j + 1].real == row[j].real
                j += 2
            else:
                j += 1


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 15, 16, 17, 24, 31, 32, 33, 47, 48, 63, 64])
def test_real_vs_oracle_and_lapack(n):
    rng = np.random.default_rng(900 + n)
    M = rng.standard_normal((150, n, n))
    w, f = eigs(M)
    assert not f.any()
    wo, io = oracle.method_dgeev(M)
    wl, il = oracle.lapack_dgeev(M)
    assert oracle.max_relerr(wo, w) <= TOL
    assert oracle.max_relerr(wl, w) <= TOL
    check_pairs(w)


def test_real_closed_forms():
    rng = np.random.default_rng(91)
    T = np.triu(rng.standard_normal((40, 40)))
    w, f = eigs(T[None])
    np.testing.assert_array_equal(np.sort(w[0].real), np.sort(np.diag(T)))
    assert not w[0].imag.any()
    ts, rs = rng.uniform(0.1, 3.0, 16), rng.uniform(0.5, 2.0, 16)
    R = np.zeros((32, 32))
    for k, (t, r) in enumerate(zip(ts, rs)):
        R[2 * k:2 * k + 2, 2 * k:2 * k + 2] = r * np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]])
    Q, _ = np.linalg.qr(rng.standard_normal((32, 32)))
    w, f = eigs((Q @ R @ Q.T)[None])
    exact = np.concatenate([rs * np.exp(1j * ts), rs * np.exp(-1j * ts)])
    assert oracle.max_relerr(exact[None], w) <= 1e-12
    check_pairs(w)
    for n in (7, 16, 50, 64):
        C = np.zeros((n, n))
        C[1:, :-1] = np.eye(n - 1)
        C[0, -1] = 1.0
        w, f = eigs(C[None])
        assert not f.any()
        assert oracle.max_relerr(np.exp(2j * np.pi * np.arange(n) / n)[None], w) <= 1e-11
    S = rng.standard_normal((60, 60))
    S = S + S.T
    w, f = eigs(S[None])
    assert np.abs(w.imag).max() == 0.0
    assert oracle.max_relerr(np.linalg.eigvalsh(S)[None].astype(complex), w) <= 1e-12
    Z = np.zeros((3, 20, 20))
    w, f = eigs(Z)
    assert not f.any() and np.abs(w).max() == 0.0


def test_real_flags():
    rng = np.random.default_rng(92)
    for n in (10, 40):
        M = rng.standard_normal((4, n, n))
        M[1, 3, 2] = np.nan
        M[2, 0, n - 1] = np.inf
        w, f = eigs(M)
        assert list(f) == [0, 2, 2, 0]
        assert np.isnan(w[1].real).all() and np.isnan(w[2].imag).all()
        assert oracle.max_relerr(oracle.lapack_dgeev(M[[0, 3]])[0], w[[0, 3]]) <= TOL
    h = tk.Handle()
    h.set_max_sweeps(0)
    w, f = eigs(rng.standard_normal((3, 30, 30)), h)
    h.close()
    assert (f == 1).all()


def test_real_full_size_invariants():
    n, B = 64, 100_000
    g = torch.Generator(device="cuda").manual_seed(17)
    M = torch.randn(B, n, n, generator=g, device="cuda", dtype=torch.float64)
    print(f"[case] real invariants n={n} batch={B}", flush=True)
    w, f = tk.eigvals_real(M)
    torch.cuda.synchronize()
    assert int((f != 0).sum()) == 0
    tr = torch.diagonal(M, dim1=-2, dim2=-1).sum(-1)
    rel = ((w.real.sum(-1) - tr).abs() / M.pow(2).sum((-1, -2)).sqrt()).max().item()
    assert rel < 1e-12 * n
    idx = torch.arange(0, B, 499)
    ws = w[idx].cpu().numpy()
    check_pairs(ws)                                          # exact conjugate pairs
    Ms = M[idx].cpu().numpy()
    # real Gaussian matrices have real eigenvalues arbitrarily close to 0,
    # where the relative error of any backward-stable method grows: judged in
    # units of the first-order bound kappa_i eps ||A||_F (as hard spectra)
    assert float(oracle.condscaled_err(Ms, ws).max()) <= 100.0
    wl, _ = oracle.lapack_dgeev(Ms)
    for b in range(len(Ms)):
        g = ws[b][oracle.match(wl[b], ws[b])]
        assert np.abs(g - wl[b]).max() <= 1e-13 * np.linalg.norm(Ms[b])


@pytest.mark.parametrize("n", [8, 17, 33, 50, 64])
def test_real_hard_spectra(n):
    M, kinds = hard_batch_real(200, n, 410 + n)
    w, f = eigs(M)
    assert not f.any()
    check_pairs(w)
    err = oracle.condscaled_err(M, w).max(axis=1)
    print(f"real hard n={n}: companion {err[0::2].max():.3g} clustered {err[1::2].max():.3g}", flush=True)
    assert float(err.max()) <= 100.0
    wo, _ = oracle.method_dgeev(M)
    assert float(oracle.condscaled_err(M, wo).max()) <= 100.0
