"""GPU checks of the ragged differentiation entries
(tokura_zgeev_ragged_schur_gpu / tokura_zgeev_ragged_backward_gpu).

* Every matrix of a ragged batch (n = 1..64, all three size classes) gets
  bit-identical eigenvalues, Schur factors and dL/dA to the same-size entries
  called on that matrix alone (same kernels' arithmetic).
* dL/dA against the oracle's adjoint (<= 1e-9 relative to max |dL/dA|).
* Hard spectra (companions and clustered, kappa(V) = 1e4), ragged fwd+bwd:
  flags 0, forward eigenvalues within the condition-scaled bound (<= 100), and
  dL/dA against the oracle within 1e-7 relative to max |dL/dA|.  The bound is
  looser here because the clustered gradients carry the eigenvector
  condition (~1/spread).
"""
import numpy as np
import pytest
import torch

import oracle
from tokura_zgeev_batched_amd.autograd import backward_batched, schur_batched
from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
from tokura_zgeev_batched_amd.workloads import hard_batch

pytestmark = pytest.mark.gpu


def rand(rng, n):
    return rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))


def oracle_grad_in_order(Mb, g, order_eigs):
    G, w = oracle.method_adjoint(Mb, np.zeros(Mb.shape[0]))
    perm = oracle.match(order_eigs, w)
    g_o = np.zeros_like(g)
    g_o[perm] = g
    G, _ = oracle.method_adjoint(Mb, g_o)
    return G


def ragged_fwd_bwd(mats, g_list):
    flat, offs = pack(mats)
    rb = RaggedBatch(offs)
    A = torch.tensor(flat, device="cuda")
    eig, schur, flag = rb.schur(A)
    gflat = torch.tensor(np.concatenate(g_list), device="cuda")
    dA = rb.backward(schur, gflat, flag)
    torch.cuda.synchronize()
    return rb, offs, eig.cpu().numpy(), schur.cpu().numpy(), flag.cpu().numpy(), dA.cpu().numpy()


def test_ragged_matches_same_size_bitwise():
    rng = np.random.default_rng(60)
    sizes = [1, 2, 5, 16, 17, 31, 32, 33, 47, 48, 49, 64] + list(rng.integers(1, 65, 40))
    mats = [rand(rng, int(n)) for n in sizes]
    g_list = [rng.standard_normal(m.shape[0]) + 1j * rng.standard_normal(m.shape[0]) for m in mats]
    print(f"[case] ragged diff bitwise, {len(mats)} matrices", flush=True)
    rb, offs, eig, schur, flag, dA = ragged_fwd_bwd(mats, g_list)
    assert not flag.any()
    for b, M in enumerate(mats):
        n = M.shape[0]
        A1 = torch.tensor(M.T.copy()[None], device="cuda")            # column-major, batch of one
        e1, s1, f1 = schur_batched(A1, n, 1)
        d1, _ = backward_batched(n, 1, torch.tensor(g_list[b][None], device="cuda"), schur=s1, flag=f1)
        torch.cuda.synchronize()
        o, eo = int(offs[b]), int(rb.eig_offsets[b])
        np.testing.assert_array_equal(eig[eo:eo + n], e1.cpu().numpy()[0])
        np.testing.assert_array_equal(schur[2 * o:2 * o + 2 * n * n], s1.cpu().numpy().reshape(-1))
        np.testing.assert_array_equal(dA[o:o + n * n], d1.cpu().numpy().reshape(-1))


def test_ragged_adjoint_vs_oracle():
    rng = np.random.default_rng(61)
    sizes = rng.integers(2, 65, 30)
    mats = [rand(rng, int(n)) for n in sizes]
    g_list = [rng.standard_normal(m.shape[0]) + 1j * rng.standard_normal(m.shape[0]) for m in mats]
    print("[case] ragged adjoint vs oracle", flush=True)
    rb, offs, eig, schur, flag, dA = ragged_fwd_bwd(mats, g_list)
    assert not flag.any()
    for b, M in enumerate(mats):
        n = M.shape[0]
        o, eo = int(offs[b]), int(rb.eig_offsets[b])
        G = dA[o:o + n * n].reshape(n, n).T                         # column-major -> M[r, c]
        Go = oracle_grad_in_order(M, g_list[b], eig[eo:eo + n])
        assert np.abs(G - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max()), (b, n)


def test_hard_spectra_ragged_fwd_bwd():
    mats, kinds = hard_batch(240, 8, 64, 62)
    rng = np.random.default_rng(63)
    g_list = [rng.standard_normal(m.shape[0]) + 1j * rng.standard_normal(m.shape[0]) for m in mats]
    print("[case] hard spectra ragged fwd+bwd 240", flush=True)
    rb, offs, eig, schur, flag, dA = ragged_fwd_bwd(mats, g_list)
    assert not flag.any()
    worst_e, worst_g = 0.0, 0.0
    for b, M in enumerate(mats):
        n = M.shape[0]
        o, eo = int(offs[b]), int(rb.eig_offsets[b])
        e = eig[eo:eo + n]
        worst_e = max(worst_e, float(oracle.condscaled_err(M, e[None]).max()))
        if b % 4 == 0:                                              # oracle adjoint on a quarter
            G = dA[o:o + n * n].reshape(n, n).T
            Go = oracle_grad_in_order(M, g_list[b], e)
            worst_g = max(worst_g, float(np.abs(G - Go).max() / max(1.0, np.abs(Go).max())))
    print(f"hard ragged fwd+bwd: condition-scaled eig err {worst_e:.3g}, grad rel diff {worst_g:.3g}", flush=True)
    assert worst_e <= 100.0
    assert worst_g <= 1e-7
