"""Pins the CPU oracle (test infrastructure) before anything is checked against it.

* golden vectors written from host LAPACK (tests/golden/make_golden.py);
* LAPACK zgehrd+zhseqr run live: the same method, implemented by OpenBLAS;
* closed forms and invariants chosen so that plausible slips (a dropped
  conjugate, a wrong shift sign, a transposed operand, a missing term in the
  adjoint) fail at least one of them.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "zgeev_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def rand(rng, b, n):
    return rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))


def test_golden_lapack_reproduces(gold):
    # the live LAPACK still gives the committed answers
    for n in gold["sizes"]:
        M = gold[f"A_{n}"]
        e, info = oracle.lapack_zgeev(M)
        assert not info.any()
        assert oracle.max_relerr(gold[f"eig_zgeev_{n}"], e) < 1e-13


def test_method_matches_golden(gold):
    for n in gold["sizes"]:
        M = gold[f"A_{n}"]
        e, flag = oracle.method_zgeev(M)
        assert not flag.any()
        assert oracle.max_relerr(gold[f"eig_zgeev_{n}"], e) < 1e-11, n
        assert oracle.max_relerr(gold[f"eig_hseqr_{n}"], e) < 1e-11, n


@pytest.mark.parametrize("n", [2, 3, 6, 11, 20, 40, 64])
def test_method_vs_lapack_same_method(n):
    rng = np.random.default_rng(n)
    M = rand(rng, 40, n)
    e, flag = oracle.method_zgeev(M)
    ref, info = oracle.lapack_hrd_hseqr(M)
    assert not flag.any() and not info.any()
    assert oracle.max_relerr(ref, e) < 1e-11


def test_triangular_and_diagonal_closed_form():
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 17, 64):
        U = np.triu(rand(rng, 1, n)[0])
        e, flag, sw = oracle.method_zgeev(U, with_sweeps=True)
        assert flag[0] == 0
        np.testing.assert_allclose(np.sort_complex(e[0]), np.sort_complex(np.diag(U)), rtol=1e-13, atol=1e-13)
        D = np.diag(np.diag(U))
        e, flag, sw = oracle.method_zgeev(D, with_sweeps=True)
        assert sw[0] == 0   # already deflated: no QR sweep
        np.testing.assert_array_equal(e[0], np.diag(D))   # order of the diagonal


def test_2x2_quadratic():
    rng = np.random.default_rng(2)
    M = rand(rng, 50, 2)
    e, _ = oracle.method_zgeev(M)
    a, b, c, d = M[:, 0, 0], M[:, 0, 1], M[:, 1, 0], M[:, 1, 1]
    disc = np.sqrt((a - d) ** 2 + 4 * b * c)
    ref = np.stack([(a + d + disc) / 2, (a + d - disc) / 2], 1)
    assert oracle.max_relerr(ref, e) < 1e-12


@pytest.mark.parametrize("n", [3, 8, 16, 33, 64])
def test_companion_roots_of_unity(n):
    C = np.zeros((n, n), complex)
    C[1:, :-1] = np.eye(n - 1)
    C[0, -1] = 1.0            # companion of z^n - 1
    e, flag = oracle.method_zgeev(C)
    assert flag[0] == 0
    ref = np.exp(2j * np.pi * np.arange(n) / n)[None]
    assert oracle.max_relerr(ref, e) < 1e-12


def test_hermitian_real_spectrum():
    rng = np.random.default_rng(3)
    X = rand(rng, 10, 24)
    Hm = X + np.conj(X.transpose(0, 2, 1))
    e, _ = oracle.method_zgeev(Hm)
    ref = np.linalg.eigvalsh(Hm)
    assert oracle.max_relerr(ref.astype(complex), e) < 1e-12
    assert np.abs(e.imag).max() < 1e-12 * np.abs(e).max()


def test_known_spectrum_similarity():
    rng = np.random.default_rng(4)
    n = 20
    lam = rand(rng, 1, n)[0][0]
    V = rand(rng, 1, n)[0]
    A = V @ np.diag(lam) @ np.linalg.inv(V)
    e, _ = oracle.method_zgeev(A)
    assert oracle.max_relerr(lam[None], e) < 1e-9


def test_conjugate_transpose_invariant():
    # eig(A^H) = conj(eig(A)): fails for a dropped conjugate in a reflector
    rng = np.random.default_rng(5)
    M = rand(rng, 20, 12)
    e1, _ = oracle.method_zgeev(M)
    e2, _ = oracle.method_zgeev(np.conj(M.transpose(0, 2, 1)))
    assert oracle.max_relerr(np.conj(e1), e2) < 1e-11


def test_matched_error_metric_detects_error():
    rng = np.random.default_rng(6)
    M = rand(rng, 4, 10)
    e, _ = oracle.method_zgeev(M)
    bad = e.copy()
    bad[2, 3] *= (1 + 1e-8)
    assert oracle.max_relerr(e, bad) > 5e-9
    perm = e[:, ::-1]
    assert oracle.max_relerr(e, perm) < 1e-15   # order does not matter


def test_hessenberg_invariants():
    rng = np.random.default_rng(7)
    for n in (2, 5, 16, 64):
        A = rand(rng, 1, n)[0]
        H = oracle.hessenberg(A)
        assert np.all(np.tril(H, -2) == 0)
        sub = np.diag(H, -1)
        assert np.all(sub.imag == 0)                      # LAPACK: real sub-diagonal
        assert abs(np.linalg.norm(H) - np.linalg.norm(A)) < 1e-13 * np.linalg.norm(A)
        assert abs(np.trace(H) - np.trace(A)) < 1e-12 * np.linalg.norm(A)


def test_schur_factorisation():
    rng = np.random.default_rng(8)
    for n in (1, 3, 16, 40):
        A = rand(rng, 1, n)[0]
        T, U, w, info = oracle.method_schur(A)
        assert info == 0
        assert np.all(np.tril(T, -1) == 0)
        assert np.linalg.norm(U.conj().T @ U - np.eye(n)) < 1e-13 * n
        assert np.linalg.norm(U @ T @ U.conj().T - A) < 1e-13 * n * np.linalg.norm(A)
        np.testing.assert_allclose(np.diag(T), w, rtol=0, atol=0)


def test_flags():
    rng = np.random.default_rng(9)
    M = rand(rng, 3, 10)
    M[1, 2, 3] = np.nan
    e, flag = oracle.method_zgeev(M)
    assert list(flag) == [0, 2, 0]
    assert np.isnan(e[1]).all()
    e, flag = oracle.method_zgeev(rand(rng, 3, 10), itmax=0)
    assert (flag == 1).all()
    e, flag = oracle.method_zgeev(np.array([[[2 + 1j]]]))
    assert flag[0] == 0 and e[0, 0] == 2 + 1j     # n = 1: converged (reference returns 1)


def _adjoint_torch(A, g):
    torch = pytest.importorskip("torch")
    At = torch.tensor(A, dtype=torch.complex128, requires_grad=True)
    lam = torch.linalg.eigvals(At)
    return lam.detach().numpy(), At, lam


def test_adjoint_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(10)
    for n in (2, 5, 12):
        A = rand(rng, 1, n)[0]
        g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        G, w = oracle.method_adjoint(A, g)
        At = torch.tensor(A, requires_grad=True)
        lam = torch.linalg.eigvals(At)
        # pair torch's order with the oracle's (Schur-diagonal) order
        perm = oracle.match(w, lam.detach().numpy())
        gt = torch.zeros(n, dtype=torch.complex128)
        gt[torch.tensor(perm)] = torch.tensor(g)
        loss = (torch.conj(gt) * lam).real.sum()
        loss.backward()
        assert np.abs(At.grad.numpy() - G).max() < 1e-10 * max(1.0, np.abs(G).max())


def test_adjoint_vs_finite_differences():
    rng = np.random.default_rng(11)
    n, h = 6, 1e-6
    A = rand(rng, 1, n)[0]
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    G, w = oracle.method_adjoint(A, g)

    def loss(X):
        e, _ = oracle.method_zgeev(X)
        perm = oracle.match(w, e[0])
        return float(np.real(np.sum(np.conj(g) * e[0][perm])))

    for (r, c) in [(0, 0), (2, 3), (5, 1)]:
        for unit in (1.0, 1j):
            E = np.zeros_like(A)
            E[r, c] = unit
            fd = (loss(A + h * E) - loss(A - h * E)) / (2 * h)
            an = np.real(np.conj(G[r, c]) * unit)   # dL = Re(conj(G) dA)
            assert abs(fd - an) < 1e-6 * max(1.0, abs(an))


def test_hard_spectra_generator_and_condition_scaled_metric():
    from tokura_zgeev_batched_amd.workloads import hard_batch, hard_spectrum, companion
    rng = np.random.default_rng(13)
    # companion: eigenvalues are the polynomial roots
    c = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    e, _ = oracle.method_zgeev(companion(c))
    roots = np.roots(np.concatenate([[1.0], c[::-1]]))
    assert oracle.max_relerr(roots[None], e) < 1e-10
    # clustered similarity: kappa(V) = 1e4 and three clusters of spread ~1e-3
    A, lam = hard_spectrum(24, rng)
    e, _ = oracle.method_zgeev(A)
    assert np.max(np.abs(e[0][oracle.match(lam, e[0])] - lam)) < 1e-6
    mats, kinds = hard_batch(60, 8, 30, 3)
    assert kinds[0] == "companion" and kinds[1] == "clustered"
    worst = 0.0
    for M in mats:
        e, f = oracle.method_zgeev(M)
        assert f[0] == 0
        worst = max(worst, oracle.condscaled_err(M, e).max())
    assert worst < 100.0
    # the metric detects an error of 1e3 condition units
    M = mats[1]
    e, _ = oracle.method_zgeev(M)
    w, rc, _ = oracle.lapack_rconde(M[None])
    bad = e.copy()
    j = int(np.argmax(rc[0]))
    i = int(np.argmin(np.abs(bad[0] - w[0][j])))
    bad[0, i] += 1e4 * np.finfo(float).eps * np.linalg.norm(M) / rc[0][j]
    assert oracle.condscaled_err(M, bad).max() > 1e3


def _check_normalised(V):
    # ZGEEV: unit 2-norm, the component of largest modulus real, >= 0 and
    # with an imaginary part of exactly zero
    for v in V.transpose(1, 0):
        assert abs(np.linalg.norm(v) - 1.0) < 1e-14
        k = int(np.argmax(np.abs(v) ** 2))
        assert v[k].imag == 0.0 and v[k].real > 0.0


@pytest.mark.parametrize("n", [1, 2, 5, 16, 33])
def test_vectors_vs_lapack_zgeev(n):
    rng = np.random.default_rng(20 + n)
    M = rand(rng, 6, n)
    w, VL, VR, info = oracle.method_vectors(M)
    wl, VLl, VRl, il = oracle.lapack_zgeev_vectors(M)
    assert not info.any() and not il.any()
    for b in range(M.shape[0]):
        perm = oracle.match(wl[b], w[b])
        assert np.abs(w[b][perm] - wl[b]).max() < 1e-12 * np.abs(wl[b]).max()
        # same normalisation => the same vectors up to the eigenvector condition
        assert np.abs(VR[b][:, perm] - VRl[b]).max() < 1e-10
        assert np.abs(VL[b][:, perm] - VLl[b]).max() < 1e-10
        nA = np.linalg.norm(M[b])
        assert np.abs(M[b] @ VR[b] - VR[b] * w[b]).max() < 1e-13 * n * nA
        assert np.abs(VL[b].conj().T @ M[b] - w[b][:, None] * VL[b].conj().T).max() < 1e-13 * n * nA
        _check_normalised(VR[b])
        _check_normalised(VL[b])


def test_vectors_closed_forms():
    # diagonal: unit vectors in diagonal order
    d = np.array([3 + 1j, -2.0, 0.5j, 7 - 2j])
    w, VL, VR, info = oracle.method_vectors(np.diag(d))
    np.testing.assert_array_equal(w[0], d)
    np.testing.assert_array_equal(VR[0], np.eye(4))
    np.testing.assert_array_equal(VL[0], np.eye(4))
    # 2x2 upper triangular [[a, b], [0, d]]: right vectors e1 and (b, d - a),
    # left vectors (d - a, -b)^* ... written out: u1 = (1, conj(b)/conj(a - d)) scaled, u2 = e2
    a, b, dd = 1 + 2j, 3 - 1j, -2 + 0.5j
    T = np.array([[a, b], [0, dd]])
    w, VL, VR, info = oracle.method_vectors(T)
    assert np.allclose(w[0], [a, dd], rtol=0, atol=1e-15)

    def zgeev_norm(v):
        v = v / np.linalg.norm(v)
        k = int(np.argmax(np.abs(v) ** 2))
        return v * (np.conj(v[k]) / abs(v[k]))

    np.testing.assert_allclose(VR[0][:, 0], [1, 0], atol=1e-15)
    np.testing.assert_allclose(VR[0][:, 1], zgeev_norm(np.array([b, dd - a])), atol=1e-15)
    np.testing.assert_allclose(VL[0][:, 1], [0, 1], atol=1e-15)
    np.testing.assert_allclose(VL[0][:, 0], zgeev_norm(np.array([1, np.conj(b) / np.conj(a - dd)])), atol=1e-15)


def test_vectors_hermitian_left_equals_right():
    rng = np.random.default_rng(23)
    X = rand(rng, 3, 12)
    Hm = X + np.conj(X.transpose(0, 2, 1))
    w, VL, VR, info = oracle.method_vectors(Hm)
    for b in range(3):
        np.testing.assert_allclose(VL[b], VR[b], atol=1e-12)
        np.testing.assert_allclose(VR[b].conj().T @ VR[b], np.eye(12), atol=1e-12)


def _graded(rng, b, n, span):
    M = rand(rng, b, n)
    d = 2.0 ** np.linspace(0, span, n)
    return (d[None, :, None] * M) / d[None, None, :], M


def _balancing_cases(rng):
    cases = []
    for n in (2, 5, 16, 40, 64):
        cases.append(rand(rng, 1, n)[0])
        cases.append(_graded(rng, 1, n, 40)[0][0])
    n = 12
    T = np.triu(rand(rng, 1, n)[0])
    T[4:9, 4:9] = rand(rng, 1, 5)[0]
    P = rng.permutation(n)
    cases.append(T[np.ix_(P, P)])                    # isolated eigenvalues after permutation
    Z = rand(rng, 1, 10)[0]
    Z[6, :] = 0
    Z[6, 6] = 2
    Z[:, 1] = 0
    Z[1, 1] = -1
    cases.append(Z)                                  # a zero row and a zero column
    return cases


def test_zgebal_matches_lapack_bitwise():
    # the restatement of ZGEBAL reproduces host LAPACK zgebal exactly
    # (permutations and power-of-two scalings are exact operations)
    rng = np.random.default_rng(30)
    for M in _balancing_cases(rng):
        for job in "NPSB":
            B1, lo1, hi1, s1, i1 = oracle.zgebal(M, job)
            B2, lo2, hi2, s2, i2 = oracle.zgebal(M, job, use_lapack=True)
            assert (lo1, hi1) == (lo2, hi2), job
            np.testing.assert_array_equal(B1, B2)
            np.testing.assert_array_equal(s1, s2)


def test_balanced_method_on_graded_matrices():
    # A = D M D^-1 with D spanning 2^40: the unbalanced method loses the
    # eigenvalues (error ~ eps ||A||), the balanced one matches LAPACK zgeev
    rng = np.random.default_rng(31)
    G, M = _graded(rng, 6, 24, 40)
    ref = np.stack([np.linalg.eigvals(m) for m in M])
    e, f = oracle.method_zgeev(G, balance="B")
    assert not f.any()
    assert oracle.max_relerr(ref, e) < 1e-12
    el, _ = oracle.lapack_zgeev(G)
    assert oracle.max_relerr(el, e) < 1e-12
    e0, _ = oracle.method_zgeev(G)
    assert oracle.max_relerr(ref, e0) > 1e-6          # balancing is what makes the difference
    # balancing never changes the spectrum of a well-scaled matrix beyond rounding
    R = rand(rng, 4, 30)
    a, _ = oracle.method_zgeev(R)
    b, _ = oracle.method_zgeev(R, balance="B")
    assert oracle.max_relerr(a, b) < 1e-12


@pytest.mark.parametrize("n", [10, 11, 17, 33, 48, 64])
def test_dual_restatement_vs_lapack_and_single_shift(n):
    """The restated two-bulge QR: the same eigenvalues as LAPACK zgeev and the
    single-shift restatement (<= 1e-10), and about 23% more single-shift
    sweeps than ZLAHQR's one shift per sweep at n = 64 (both shifts of a
    sweep pair come from one trailing 2x2)."""
    rng = np.random.default_rng(1200 + n)
    M = rng.standard_normal((40, n, n)) + 1j * rng.standard_normal((40, n, n))
    wd, fd, sd = oracle.method_zgeev_dual(M, with_sweeps=True)
    ws, fs, ss = oracle.method_zgeev(M, with_sweeps=True)
    wl, _ = oracle.lapack_zgeev(M)
    assert not fd.any() and not fs.any()
    assert oracle.max_relerr(wl, wd) <= 1e-10
    assert oracle.max_relerr(ws, wd) <= 1e-10
    if n == 64:
        assert 1.1 < sd.mean() / ss.mean() < 1.4


def test_dual_restatement_closed_forms():
    # companion of z^n - 1 (needs the exceptional shifts), triangular input
    # (no sweep), a block-diagonal matrix (independent blocks after deflation)
    for n in (12, 40, 64):
        C = np.zeros((n, n), complex)
        C[1:, :-1] = np.eye(n - 1)
        C[0, -1] = 1.0
        w, f = oracle.method_zgeev_dual(C[None])
        assert not f.any()
        assert oracle.max_relerr(np.exp(2j * np.pi * np.arange(n) / n)[None], w) <= 1e-11
    rng = np.random.default_rng(77)
    T = np.triu(rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30)))
    w, f, sw = oracle.method_zgeev_dual(T[None], with_sweeps=True)
    assert sw[0] == 0
    np.testing.assert_array_equal(np.sort_complex(w[0]), np.sort_complex(np.diag(T)))
    B = np.zeros((40, 40), complex)
    B[:25, :25] = rng.standard_normal((25, 25)) + 1j * rng.standard_normal((25, 25))
    B[25:, 25:] = rng.standard_normal((15, 15)) + 1j * rng.standard_normal((15, 15))
    w, f = oracle.method_zgeev_dual(B[None])
    ref = np.concatenate([np.linalg.eigvals(B[:25, :25]), np.linalg.eigvals(B[25:, 25:])])
    assert oracle.max_relerr(ref[None], w) <= 1e-10
    w, f = oracle.method_zgeev_dual(rng.standard_normal((2, 30, 30))[None][0].astype(complex), itmax=0)
    assert (f == 1).all()


def test_norm_range_scaling_vs_lapack():
    """ZGEEV's norm-range scaling in the restatement: at entries near 1e-300 the
    QR's deflation threshold (safmin n / ulp, about 1e-291) is no longer far
    below the sub-diagonal, and products of two entries underflow; LAPACK
    zgeev scales A into [2^-459, 2^459] first.  The restatement (single- and
    two-bulge QR) must match LAPACK zgeev at every scale, and its power-of-two
    factor must leave in-range matrices untouched: eig(2^k A) = 2^k eig(A)
    bit for bit while 2^k A stays in range."""
    rng = np.random.default_rng(4)
    M = rng.standard_normal((12, 24, 24)) + 1j * rng.standard_normal((12, 24, 24))
    base, _ = oracle.method_zgeev(M)
    for s in (1e-305, 1e-300, 1e-290, 1e-250, 1e-200, 1e-139, 1e139, 1e200, 1e250, 1e290, 1e300):
        ref, _ = oracle.lapack_zgeev(M * s)
        for fn in (oracle.method_zgeev, oracle.method_zgeev_dual):
            e, f = fn(M * s)
            assert not f.any()
            assert oracle.max_relerr(ref, e) <= 1e-12, (s, fn.__name__)
    for k in (-400, -100, 100, 400):
        e, _ = oracle.method_zgeev(M * 2.0 ** k)
        np.testing.assert_array_equal(e, base * 2.0 ** k)
