"""GPU checks of the eigenvector entry (tokura_zgeev_batched_vectors_gpu).

* against the oracle's restatement of the same method (Schur form, ZTREVC,
  back-transform, ZGEEV normalisation) and against host LAPACK zgeev('V','V'):
  the same normalisation makes the vectors directly comparable;
* residuals ||A V - V diag(w)|| and ||VL^H A - diag(w) VL^H|| relative to
  ||A||, and the ZGEEV normalisation (unit norm, largest component real);
* eigenvalues identical to the Schur forward's, in the same order;
* left-only / right-only calls, n = 1, and the NaN flag.
Bar: eigenvalues <= 1e-10 relative vs LAPACK; eigenvectors <= 1e-10 vs the
oracle and <= 1e-8 vs LAPACK (balanced, so differs by the eigenvector
condition); residuals <= 1e-13 * n * ||A||.
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.autograd import schur_batched

pytestmark = pytest.mark.gpu


def rand(rng, b, n):
    return rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))


def normalised(V):
    for v in V.T:
        assert abs(np.linalg.norm(v) - 1.0) < 1e-13
        k = int(np.argmax(np.abs(v) ** 2))
        assert v[k].imag == 0.0 and v[k].real > 0.0


@pytest.mark.parametrize("n", [1, 2, 7, 16, 17, 31, 32, 33, 48, 64])
def test_vectors_vs_oracle_and_lapack(gpu, n):
    rng = np.random.default_rng(700 + n)
    B = 24
    M = rand(rng, B, n)
    print(f"[case] vectors n={n} B={B}", flush=True)
    w, VL, VR, flag = tk.eig(torch.tensor(M, device=gpu), left=True, right=True)
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    w, VL, VR = w.cpu().numpy(), VL.cpu().numpy(), VR.cpu().numpy()
    wo, VLo, VRo, info = oracle.method_vectors(M)
    wl, VLl, VRl, il = oracle.lapack_zgeev_vectors(M)
    assert not info.any() and not il.any()
    assert oracle.max_relerr(wl, w) <= 1e-10
    for b in range(B):
        nA = np.linalg.norm(M[b])
        assert np.abs(M[b] @ VR[b] - VR[b] * w[b]).max() <= 1e-13 * n * nA
        assert np.abs(VL[b].conj().T @ M[b] - w[b][:, None] * VL[b].conj().T).max() <= 1e-13 * n * nA
        normalised(VR[b])
        normalised(VL[b])
        po = oracle.match(wo[b], w[b])      # w[b][po] pairs with the oracle's order
        assert np.abs(VR[b][:, po] - VRo[b]).max() <= 1e-10
        assert np.abs(VL[b][:, po] - VLo[b]).max() <= 1e-10
        pl = oracle.match(wl[b], w[b])
        assert np.abs(VR[b][:, pl] - VRl[b]).max() <= 1e-8
        assert np.abs(VL[b][:, pl] - VLl[b]).max() <= 1e-8


@pytest.mark.parametrize("n", [5, 20, 32, 40, 57, 64])
def test_vectors_one_side_and_schur_order(gpu, n):
    rng = np.random.default_rng(800 + n)
    M = torch.tensor(rand(rng, 16, n), device=gpu)
    print(f"[case] vectors one side n={n}", flush=True)
    w_lr, VL, VR, f0 = tk.eig(M, left=True, right=True)
    w_r, _, VR_only, f1 = tk.eig(M, left=False, right=True)                    # T parked in the workspace
    w_l, VL_only, _, f2 = tk.eig(M, left=True, right=False)
    w_r1, _, VR_1k, f4 = tk.eig(M, left=False, right=True, workspace=False)    # one-kernel form
    w_l1, VL_1k, _, f5 = tk.eig(M, left=True, right=False, workspace=False)
    A = M.transpose(-1, -2).contiguous()
    e_s, _, f3 = schur_batched(A, n, 16)
    torch.cuda.synchronize()
    for f in (f0, f1, f2, f3, f4, f5):
        assert int(f.abs().sum()) == 0
    # one Schur form behind every call: identical eigenvalues and vectors
    for w in (w_lr, w_r, w_l, w_r1, w_l1):
        assert torch.equal(w, e_s)
    assert torch.equal(VR, VR_only) and torch.equal(VL, VL_only)
    assert torch.equal(VR, VR_1k) and torch.equal(VL, VL_1k)


def test_vectors_closed_forms(gpu):
    d = np.array([3 + 1j, -2.0, 0.5j, 7 - 2j, 1.0])
    w, VL, VR, flag = tk.eig(torch.tensor(np.diag(d)[None], device=gpu), left=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(w.cpu().numpy()[0], d)
    np.testing.assert_array_equal(VR.cpu().numpy()[0], np.eye(5))
    np.testing.assert_array_equal(VL.cpu().numpy()[0], np.eye(5))
    w, VL, VR, flag = tk.eig(torch.tensor([[[2.0 - 1j]]], dtype=torch.complex128, device=gpu), left=True)
    torch.cuda.synchronize()
    assert int(flag[0]) == 0 and complex(w[0, 0]) == 2 - 1j
    assert complex(VR[0, 0, 0]) == 1 and complex(VL[0, 0, 0]) == 1


def test_vectors_nan_flag(gpu):
    rng = np.random.default_rng(9)
    M = rand(rng, 3, 12)
    M[1, 4, 5] = np.nan
    w, VL, VR, flag = tk.eig(torch.tensor(M, device=gpu), left=True)
    torch.cuda.synchronize()
    assert flag.cpu().tolist() == [0, 2, 0]
    assert torch.isnan(VR[1].real).all() and torch.isnan(VL[1].real).all() and torch.isnan(w[1].real).all()
    assert not torch.isnan(VR[0].real).any() and not torch.isnan(VR[2].real).any()
    for M2 in (M, rand(rng, 3, 40)):          # one-sided, both size classes of the split path
        M2[1, 4, 5] = np.nan
        w, VL, VR, flag = tk.eig(torch.tensor(M2, device=gpu), left=False, right=True)
        torch.cuda.synchronize()
        assert flag.cpu().tolist() == [0, 2, 0] and VL is None
        assert torch.isnan(VR[1].real).all() and not torch.isnan(VR[[0, 2]].real).any()
