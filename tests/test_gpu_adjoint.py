"""GPU checks of the differentiation path through the C ABI.

* forward with Schur vectors: A = U T U^H, U unitary, T upper triangular,
  diag(T) = eigenvalues (<= 1e-10 vs LAPACK zgeev), same values in the same
  order as the matching eigenvalue-only variant (n <= 32: single-bulge QR,
  n > 32: two-bulge QR);
* adjoint: vs the oracle's dL/dA (same method on the CPU), vs CPU
  torch.linalg.eigvals autograd in complex128, and vs central finite
  differences (h = 1e-6) with gradient relative error <= 1e-7;
* recompute backward (tokura_zgeev_batched_backward_recompute_gpu): its
  pairing of the forward's eigenvalues with its own Schur order, after the
  default (autotuned) forward, after a shuffled forward, and a forward that
  belongs to other matrices (flag 3).
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.autograd import backward_batched, backward_recompute, eigvals_autograd, schur_batched

pytestmark = pytest.mark.gpu


def rand(rng, b, n):
    return rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))


@pytest.mark.parametrize("n", [1, 2, 5, 16, 17, 18, 32, 33, 40, 64])
def test_schur_forward(n):
    rng = np.random.default_rng(300 + n)
    M = rand(rng, 20, n)
    A = torch.tensor(M, device="cuda").transpose(-1, -2).contiguous()
    print(f"[case] schur n={n}", flush=True)
    eig, schur, flag = schur_batched(A, n, 20)
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    T = schur[:, 0].transpose(-1, -2).cpu().numpy()
    U = schur[:, 1].transpose(-1, -2).cpu().numpy()
    e = eig.cpu().numpy()
    for b in range(20):
        assert np.all(np.tril(T[b], -1) == 0)
        assert np.linalg.norm(U[b].conj().T @ U[b] - np.eye(n)) < 1e-13 * n
        assert np.linalg.norm(U[b] @ T[b] @ U[b].conj().T - M[b]) < 1e-13 * n * np.linalg.norm(M[b])
        np.testing.assert_array_equal(np.diag(T[b]), e[b])
    ref, _ = oracle.lapack_zgeev(M)
    assert oracle.max_relerr(ref, e) <= 1e-10


def _colmajor(M):
    return torch.tensor(M, device="cuda").transpose(-1, -2).contiguous()


@pytest.mark.parametrize("n", [1, 3, 5, 16, 17, 24, 32, 33, 40, 48, 57, 64])
def test_recompute_after_default_forward_vs_oracle(n):
    """The eigenvalue forward with the handle's default variants (whose order
    need not be the Schur form's), then the recompute backward: dL/dA equals
    the oracle's with the cotangents in the forward's order."""
    rng = np.random.default_rng(700 + n)
    B = 8
    M = rand(rng, B, n)
    g = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    A = _colmajor(M)
    print(f"[case] default forward + recompute backward n={n}", flush=True)
    eig, flag = tk.eigvals(torch.tensor(M, device="cuda"))
    dA, f2 = backward_recompute(n, B, A, eig, torch.tensor(g, device="cuda"))
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0 and int(f2.abs().sum()) == 0
    G = dA.transpose(-1, -2).cpu().numpy()
    e = eig.cpu().numpy()
    for b in range(B):
        Go = oracle_grad_in_order(M[b], g[b], e[b])
        assert np.abs(G[b] - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max()), b


@pytest.mark.parametrize("n", [2, 9, 16, 20, 32, 36, 50, 64])
def test_recompute_pairs_shuffled_forward(n):
    """Forward eigenvalues and cotangents shuffled by the same random
    permutation per matrix give the same dL/dA bits as the recompute backward
    with them in the Schur order (the snapped shifts do not depend on the
    order of the known eigenvalues), and agree with the stored-Schur backward
    to 1e-9 (the recompute's Schur form comes from other shifts)."""
    rng = np.random.default_rng(800 + n)
    B = 12
    M = rand(rng, B, n)
    A = _colmajor(M)
    g = torch.tensor(rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n)), device="cuda")
    e_s, schur, fl = schur_batched(A, n, B)
    dA_store, _ = backward_batched(n, B, g, schur, fl.clone())
    perm = torch.tensor(np.stack([rng.permutation(n) for _ in range(B)]), device="cuda")
    e_p = torch.gather(e_s, 1, perm)
    g_p = torch.gather(g, 1, perm)
    print(f"[case] shuffled forward n={n}", flush=True)
    dA_rc, f2 = backward_recompute(n, B, A, e_p, g_p)
    dA_rs, f3 = backward_recompute(n, B, A, e_s, g)
    torch.cuda.synchronize()
    assert int(f2.abs().sum()) == 0 and int(f3.abs().sum()) == 0
    assert torch.equal(dA_rc, dA_rs)
    scale = dA_store.abs().amax(dim=(1, 2)).clamp(min=1.0)
    assert float(((dA_rc - dA_store).abs().amax(dim=(1, 2)) / scale).max()) <= 1e-9


@pytest.mark.parametrize("n", [4, 16, 30, 64])
def test_recompute_flags_foreign_eigenvalues(n):
    """Eigenvalues of other matrices: flag 3 and dA = NaN for those
    matrices only."""
    rng = np.random.default_rng(850 + n)
    B = 6
    M = rand(rng, B, n)
    A = _colmajor(M)
    eig, _ = tk.eigvals(torch.tensor(M, device="cuda"))
    bad = eig.clone()
    bad[1] = eig[2]
    bad[4] = eig[4] * 3.0
    g = torch.ones((B, n), dtype=torch.complex128, device="cuda")
    dA, f = backward_recompute(n, B, A, bad, g)
    torch.cuda.synchronize()
    assert f.cpu().numpy().tolist() == [0, 3, 0, 0, 3, 0]
    G = dA.cpu().numpy()
    assert np.isnan(G[[1, 4]]).all() and np.isfinite(G[[0, 2, 3, 5]]).all()


def test_recompute_entry_validation():
    h = tk.default_handle()
    f = tk.lib().tokura_zgeev_batched_backward_recompute_gpu
    x = torch.zeros(64, dtype=torch.complex128, device="cuda")
    p = x.data_ptr()
    assert f(None, 4, 1, p, p, p, p, p, p, None) == -1
    assert f(h.ptr, 65, 1, p, p, p, p, p, p, None) == -2
    assert f(h.ptr, 4, 1, None, p, p, p, p, p, None) == -3
    assert f(h.ptr, 4, 1, p, None, p, p, p, p, None) == -4
    assert f(h.ptr, 4, 1, p, p, p, p, None, p, None) == -5
    assert f(h.ptr, 4, 0, p, p, p, p, p, p, None) == -6
    assert f(h.ptr, 4, 1, p, p, p, p, p, None, None) == -7
    # the stored-Schur entry needs the Schur form
    assert tk.lib().tokura_zgeev_batched_backward_gpu(h.ptr, 4, 1, p, None, p, p, p, None) == -3


def oracle_grad_in_order(Mb, g, order_eigs):
    """Oracle dL/dA with g given in the order of `order_eigs`."""
    G, w = oracle.method_adjoint(Mb, np.zeros(Mb.shape[0]))
    perm = oracle.match(order_eigs, w)   # w[perm] pairs with order_eigs
    g_o = np.zeros_like(g)
    g_o[perm] = g
    G, w2 = oracle.method_adjoint(Mb, g_o)
    return G


@pytest.mark.parametrize("mode", ["store", "recompute"])
@pytest.mark.parametrize("n", [2, 7, 16, 32, 45, 64])
def test_adjoint_vs_oracle_and_torch(n, mode):
    rng = np.random.default_rng(400 + n)
    B = 6
    M = rand(rng, B, n)
    g = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    Mt = torch.tensor(M, device="cuda", requires_grad=True)
    print(f"[case] adjoint n={n} mode={mode}", flush=True)
    eig, flag = eigvals_autograd(Mt, mode=mode)
    loss = (torch.conj(torch.tensor(g, device="cuda")) * eig).real.sum()
    loss.backward()
    G = Mt.grad.cpu().numpy()
    e = eig.detach().cpu().numpy()
    for b in range(B):
        Go = oracle_grad_in_order(M[b], g[b], e[b])
        scale = max(1.0, np.abs(Go).max())
        assert np.abs(G[b] - Go).max() <= 1e-9 * scale, b
        # CPU torch autograd, complex128
        At = torch.tensor(M[b], requires_grad=True)
        lam = torch.linalg.eigvals(At)
        perm = oracle.match(e[b], lam.detach().numpy())   # lam[perm] pairs with e
        gt = torch.zeros(n, dtype=torch.complex128)
        gt[torch.tensor(perm)] = torch.tensor(g[b])
        (torch.conj(gt) * lam).real.sum().backward()
        assert np.abs(At.grad.numpy() - G[b]).max() <= 1e-8 * scale, b


def _loss_fd(M, g, e0, E, h, handle=None):
    """L(M + h E) per matrix, eigenvalues paired with e0 by optimal matching."""
    ep, _ = tk.eigvals(torch.tensor(M + h * E, device="cuda"), handle)
    ep = ep.cpu().numpy()
    out = np.zeros(M.shape[0])
    for b in range(M.shape[0]):
        pp = oracle.match(e0[b], ep[b])
        out[b] = np.real(np.sum(np.conj(g[b]) * ep[b][pp]))
    return out


def test_adjoint_vs_finite_differences():
    # 100 sampled matrices of the n = 32 configuration; for each a random unit
    # direction E.  Gradient relative error = |FD - <G, E>| / ||G||_F, with FD
    # the central difference at h = 1e-6 through the eigenvalue-only kernel.
    rng = np.random.default_rng(11)
    n, B, h = 32, 100, 1e-6
    M = rand(rng, B, n)
    g = rng.standard_normal((B, n)) + 1j * rng.standard_normal((B, n))
    Mt = torch.tensor(M, device="cuda", requires_grad=True)
    eig, flag = eigvals_autograd(Mt)
    (torch.conj(torch.tensor(g, device="cuda")) * eig).real.sum().backward()
    G = Mt.grad.cpu().numpy()
    e0 = eig.detach().cpu().numpy()
    E = rand(rng, B, n)
    E /= np.linalg.norm(E, axis=(1, 2), keepdims=True)
    an = np.real(np.sum(np.conj(G) * E, axis=(1, 2)))
    gnorm = np.linalg.norm(G, axis=(1, 2))
    fd2 = (_loss_fd(M, g, e0, E, h) - _loss_fd(M, g, e0, E, -h)) / (2 * h)
    err2 = np.abs(fd2 - an) / gnorm
    # fourth-order stencil at a larger step: truncation O(h^4), less round-off
    h4 = 1e-3
    fd4 = (-_loss_fd(M, g, e0, E, 2 * h4) + 8 * _loss_fd(M, g, e0, E, h4)
           - 8 * _loss_fd(M, g, e0, E, -h4) + _loss_fd(M, g, e0, E, -2 * h4)) / (12 * h4)
    err4 = np.abs(fd4 - an) / gnorm
    print(f"FD gradient rel err: central h=1e-6 max {err2.max():.2e} median {np.median(err2):.2e}; "
          f"4th-order h=1e-3 max {err4.max():.2e} median {np.median(err4):.2e}", flush=True)
    assert err2.max() <= 1e-7
    assert err4.max() <= 1e-6


@pytest.mark.parametrize("n", [5, 17, 32, 33, 40, 47, 50, 64])
def test_store_and_recompute_gradients_agree(n):
    """The stored-Schur backward and the recompute backward (the two-launch
    Schur form rebuilt with T in the workspace and U in dA, its QR shifts
    snapped to the given eigenvalues, then the same adjoint arithmetic) agree
    to 1e-9 of the gradient's largest entry, and the recompute gradient
    matches the oracle's to the same bar."""
    rng = np.random.default_rng(900 + n)
    M = rand(rng, 40, n)
    A = _colmajor(M)
    g = torch.tensor(rng.standard_normal((40, n)) + 1j * rng.standard_normal((40, n)), device="cuda")
    print(f"[case] store vs recompute gradients n={n}", flush=True)
    e_s, schur, fl = schur_batched(A, n, 40)
    assert int(fl.abs().sum()) == 0
    d1, _ = backward_batched(n, 40, g, schur, fl.clone())
    d2, f2 = backward_recompute(n, 40, A, e_s, g)
    torch.cuda.synchronize()
    assert int(f2.abs().sum()) == 0
    scale = d1.abs().amax(dim=(1, 2)).clamp(min=1.0)
    assert float(((d1 - d2).abs().amax(dim=(1, 2)) / scale).max()) <= 1e-9
    G = d2.transpose(-1, -2).cpu().numpy()
    e = e_s.cpu().numpy()
    gh = g.cpu().numpy()
    for b in range(0, 40, 8):
        Go = oracle_grad_in_order(M[b], gh[b], e[b])
        assert np.abs(G[b] - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max()), b


def test_adjoint_flags_nan():
    rng = np.random.default_rng(12)
    M = rand(rng, 3, 9)
    M[1, 4, 4] = np.nan
    Mt = torch.tensor(M, device="cuda", requires_grad=True)
    eig, flag = eigvals_autograd(Mt)
    assert list(flag.cpu().numpy()) == [0, 2, 0]
    (eig[[0, 2]].real.sum()).backward()
    G = Mt.grad.cpu().numpy()
    assert np.isnan(G[1]).all() and np.isfinite(G[[0, 2]]).all()


def test_diff_config_sample_full_size():
    # 50,000 matrices n = 32: forward with Schur vectors + adjoint; a sample
    # checked against the oracle
    n, B = 32, 50000
    gen = torch.Generator(device="cuda").manual_seed(3)
    M = torch.complex(torch.randn(B, n, n, generator=gen, device="cuda", dtype=torch.float64),
                      torch.randn(B, n, n, generator=gen, device="cuda", dtype=torch.float64))
    gv = torch.complex(torch.randn(B, n, generator=gen, device="cuda", dtype=torch.float64),
                       torch.randn(B, n, generator=gen, device="cuda", dtype=torch.float64))
    M.requires_grad_(True)
    print("[case] diff config 50000 x 32", flush=True)
    eig, flag = eigvals_autograd(M)
    (torch.conj(gv) * eig).real.sum().backward()
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    idx = list(range(0, B, 2500))
    Mh = M.detach()[idx].cpu().numpy()
    Gh = M.grad[idx].cpu().numpy()
    eh = eig.detach()[idx].cpu().numpy()
    gh = gv[idx].cpu().numpy()
    for k in range(len(idx)):
        Go = oracle_grad_in_order(Mh[k], gh[k], eh[k])
        assert np.abs(Gh[k] - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max())
    # recompute mode on the whole batch: the default forward, then the
    # recompute backward; a sample against the oracle
    Md = M.detach()
    A = Md.transpose(-1, -2).contiguous()
    e_d, f_d = tk.eigvals(Md)
    dA, f2 = backward_recompute(n, B, A, e_d, gv)
    torch.cuda.synchronize()
    assert int(f_d.abs().sum()) == 0 and int(f2.abs().sum()) == 0
    Gr = dA.transpose(-1, -2)[idx].cpu().numpy()
    er = e_d[idx].cpu().numpy()
    for k in range(len(idx)):
        Go = oracle_grad_in_order(Mh[k], gh[k], er[k])
        assert np.abs(Gr[k] - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max())
