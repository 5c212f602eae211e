"""GPU parity of the same-size path (tokura_zgeev_batched_gpu) through the C ABI.

Bar: max relative error <= 1e-10 against host LAPACK zgeev after optimal
eigenvalue matching (the accuracy target), and <= 1e-11 against the oracle's
restatement of the same method (zgehd2 + zlahqr without balancing).
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk

pytestmark = pytest.mark.gpu

TOL_LAPACK = 1e-10
TOL_METHOD = 1e-11


def rand(rng, b, n):
    return rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n))


def gpu_eigs(M, handle):
    print(f"[case] n={M.shape[-1]} batch={M.shape[0]}", flush=True)
    eig, flag = tk.eigvals(torch.tensor(M, device="cuda"), handle)
    torch.cuda.synchronize()
    return eig.cpu().numpy(), flag.cpu().numpy()


@pytest.mark.parametrize("n", list(range(1, 65)))
def test_every_size_default_config(handle, n):
    rng = np.random.default_rng(1000 + n)
    M = rand(rng, 33, n)
    e, flag = gpu_eigs(M, handle)
    assert not flag.any()
    ref, info = oracle.lapack_zgeev(M)
    assert oracle.max_relerr(ref, e) <= TOL_LAPACK
    me, mflag = oracle.method_zgeev(M)
    assert oracle.max_relerr(me, e) <= TOL_METHOD


@pytest.mark.parametrize("n", [1, 3, 4, 5, 8, 9, 16, 17, 31, 32, 33, 48, 49, 64])
def test_every_variant(n):
    rng = np.random.default_rng(2000 + n)
    M = rand(rng, 70, n)             # 70: a partly filled last workgroup for every variant
    ref, _ = oracle.lapack_zgeev(M)
    vs = []
    for v in range(40):
        h = tk.Handle()
        if tk.lib().tokuraSetVariant(h.ptr, n, v) != 0:
            h.close()
            continue
        vs.append(v)
        print(f"[case] variant={v}", flush=True)
        e, flag = gpu_eigs(M, h)
        h.close()
        assert not flag.any(), v
        assert oracle.max_relerr(ref, e) <= TOL_LAPACK, v
    assert vs


def test_golden(handle):
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "zgeev_golden.npz"))
    for n in g["sizes"]:
        e, flag = gpu_eigs(g[f"A_{n}"], handle)
        assert not flag.any()
        assert oracle.max_relerr(g[f"eig_zgeev_{n}"], e) <= TOL_LAPACK


def test_closed_forms(handle):
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 32, 64):
        U = np.triu(rand(rng, 1, n)[0])
        e, flag = gpu_eigs(U[None], handle)
        assert flag[0] == 0
        np.testing.assert_allclose(np.sort_complex(e[0]), np.sort_complex(np.diag(U)), rtol=1e-13, atol=1e-13)
        D = np.diag(np.diag(U))
        e, _ = gpu_eigs(D[None], handle)
        np.testing.assert_array_equal(e[0], np.diag(D))
    for n in (3, 16, 64):
        C = np.zeros((n, n), complex)
        C[1:, :-1] = np.eye(n - 1)
        C[0, -1] = 1.0
        e, flag = gpu_eigs(C[None], handle)
        assert flag[0] == 0
        assert oracle.max_relerr(np.exp(2j * np.pi * np.arange(n) / n)[None], e) < 1e-12
    X = rand(rng, 8, 40)
    Hm = X + np.conj(X.transpose(0, 2, 1))
    e, _ = gpu_eigs(Hm, handle)
    assert oracle.max_relerr(np.linalg.eigvalsh(Hm).astype(complex), e) < 1e-12
    Z = np.zeros((2, 9, 9), complex)
    e, flag = gpu_eigs(Z, handle)
    assert not flag.any() and np.all(e == 0)


@pytest.mark.parametrize("n", [33, 40, 47, 49, 55, 63])
def test_structured_inputs_offset_hessenberg(handle, n):
    """Sizes between multiples of 16, where the register Hessenberg places the
    matrix at the bottom right of its grid (zero rows and columns above and
    left of it): structured inputs whose reflectors are trivial or whose
    spectrum is known, against closed forms and host LAPACK, through the
    same-size path and the ragged path."""
    rng = np.random.default_rng(5000 + n)
    U = np.triu(rand(rng, 1, n)[0])
    D = np.diag(np.diag(U))
    Z = np.zeros((n, n), complex)
    R = rand(rng, 1, n)[0]
    R[:, n // 3] = 0.0                      # a zero column: reducible, eigenvalue 0
    # (a zero row as well would make 0 defective, its error ~sqrt(eps) ||A||)
    Hs = np.triu(rand(rng, 1, n)[0], -1)    # already upper Hessenberg
    P = np.zeros((n, n), complex)           # cyclic permutation: roots of unity
    P[1:, :-1] = np.eye(n - 1)
    P[0, -1] = 1.0
    M = np.stack([U, D, Z, R, Hs, P])
    ref_hs, _ = oracle.lapack_zgeev(M[4:5])

    def check(e):
        np.testing.assert_allclose(np.sort_complex(e[0]), np.sort_complex(np.diag(U)), rtol=1e-13, atol=1e-13)
        np.testing.assert_array_equal(e[1], np.diag(D))
        assert np.all(e[2] == 0)
        # eigenvalue 0 of R comes back as a rounding-size number: errors in
        # units of kappa_i eps ||A||_F (as the hard-spectra tests)
        assert float(oracle.condscaled_err(M[3:4], e[3:4]).max()) <= 100.0
        assert oracle.max_relerr(ref_hs, e[4:5]) <= TOL_LAPACK
        assert oracle.max_relerr(np.exp(2j * np.pi * np.arange(n) / n)[None], e[5:6]) < 1e-12

    e, flag = gpu_eigs(M, handle)
    assert not flag.any()
    check(e)
    from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
    mats = [M[i] for i in range(len(M))] + [rand(rng, 1, k)[0] for k in (5, 20, 64)]
    flat, offs = pack(mats)
    rb = RaggedBatch(offs, handle=tk.Handle())
    er, fr = rb.solve(torch.tensor(flat, device="cuda"))
    torch.cuda.synchronize()
    assert int((fr != 0).sum()) == 0
    er = er.cpu().numpy()
    check(np.stack([er[rb.eig_offsets[i]:rb.eig_offsets[i + 1]] for i in range(len(M))]))


SCALES = (1e-305, 1e-300, 1e-290, 1e-250, 1e-200, 1e-150, 1e-120, 1e-30, 1e30, 1e120, 1e150, 1e200, 1e250,
          1e290, 1e300)


@pytest.mark.parametrize("n", [4, 12, 24, 40, 57, 64])
def test_scaled_inputs(handle, n):
    """ZGEEV's norm-range scaling: entries from 1e-305 to 1e300 against host
    LAPACK zgeev on the same matrices, every default variant's size range
    (fused, split, two-bulge, pool)."""
    rng = np.random.default_rng(4 + n)
    M = rand(rng, 16, n)
    for s in SCALES:
        ref, _ = oracle.lapack_zgeev(M * s)
        e, flag = gpu_eigs(M * s, handle)
        assert not flag.any(), s
        assert oracle.max_relerr(ref, e) <= TOL_LAPACK, s
        me, _ = oracle.method_zgeev(M * s)
        assert oracle.max_relerr(me, e) <= TOL_METHOD, s


def test_scaled_inputs_every_variant():
    rng = np.random.default_rng(41)
    for n in (9, 33, 64):
        M = rand(rng, 70, n)
        for s in (1e-300, 1e-290, 1e290):
            ref, _ = oracle.lapack_zgeev(M * s)
            for v in range(40):
                h = tk.Handle()
                if tk.lib().tokuraSetVariant(h.ptr, n, v) != 0:
                    h.close()
                    continue
                e, flag = gpu_eigs(M * s, h)
                h.close()
                assert not flag.any(), (n, s, v)
                assert oracle.max_relerr(ref, e) <= TOL_LAPACK, (n, s, v)


def test_scaled_inputs_ragged():
    from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
    rng = np.random.default_rng(42)
    sizes = rng.integers(1, 65, 120)
    scales = np.array(SCALES)[rng.integers(0, len(SCALES), 120)]
    mats = [rand(rng, 1, int(n))[0] * s for n, s in zip(sizes, scales)]
    flat, offs = pack(mats)
    rb = RaggedBatch(offs)
    eig, flag = rb.solve(torch.tensor(flat, device="cuda"))
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    e = eig.cpu().numpy()
    for b, m in enumerate(mats):
        ref, _ = oracle.lapack_zgeev(m[None])
        got = e[None, rb.eig_offsets[b]:rb.eig_offsets[b + 1]]
        assert oracle.max_relerr(ref, got) <= TOL_LAPACK, (b, m.shape[0], scales[b])


def test_flags(handle):
    rng = np.random.default_rng(5)
    M = rand(rng, 4, 12)
    M[2, 5, 1] = np.inf
    M[3, 0, 0] = np.nan
    e, flag = gpu_eigs(M, handle)
    assert list(flag) == [0, 0, 2, 2]
    assert np.isnan(e[2]).all() and np.isnan(e[3]).all()
    ref, _ = oracle.lapack_zgeev(M[:2])
    assert oracle.max_relerr(ref, e[:2]) <= TOL_LAPACK
    h = tk.Handle()
    h.set_max_sweeps(0)
    e, flag = gpu_eigs(rand(rng, 5, 12), h)
    assert (flag == 1).all()
    e, flag = gpu_eigs(np.array([[[3 - 2j]]]), h)   # n = 1 needs no sweep
    assert flag[0] == 0 and e[0, 0] == 3 - 2j
    h.close()


def test_small_config_full_size(handle):
    # 10,000 matrices n = 16, every one checked against host LAPACK
    rng = np.random.default_rng(6)
    M = rand(rng, 10000, 16)
    e, flag = gpu_eigs(M, handle)
    assert not flag.any()
    ref, info = oracle.lapack_zgeev(M)
    assert oracle.max_relerr(ref, e) <= TOL_LAPACK


def test_n64_batch_trace_and_sample(handle):
    # size-independent invariants on a large n = 64 batch: sum(lambda) = tr(A),
    # sum(lambda^2) = tr(A^2); plus LAPACK on a sample
    g = torch.Generator(device="cuda").manual_seed(7)
    B, n = 20000, 64
    A = torch.complex(torch.randn(B, n, n, generator=g, device="cuda", dtype=torch.float64),
                      torch.randn(B, n, n, generator=g, device="cuda", dtype=torch.float64))
    print("[case] n=64 batch=20000 invariants", flush=True)
    eig, flag = tk.eigvals(A, handle)
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    tr = torch.diagonal(A, dim1=-2, dim2=-1).sum(-1)
    tr2 = torch.diagonal(A @ A, dim1=-2, dim2=-1).sum(-1)
    nrm = torch.linalg.matrix_norm(A)
    assert float(((eig.sum(-1) - tr).abs() / nrm).max()) < 1e-12
    assert float(((eig.pow(2).sum(-1) - tr2).abs() / nrm ** 2).max()) < 1e-12
    idx = torch.arange(0, B, 97, device="cuda")
    Ms = A[idx].cpu().numpy()
    ref, _ = oracle.lapack_zgeev(Ms)
    assert oracle.max_relerr(ref, eig[idx].cpu().numpy()) <= TOL_LAPACK


def test_stream_and_reuse(handle):
    rng = np.random.default_rng(8)
    M = rand(rng, 300, 20)
    s = torch.cuda.Stream()
    A = torch.tensor(M, device="cuda").transpose(-1, -2).contiguous()
    eig = torch.empty((300, 20), dtype=torch.complex128, device="cuda")
    flag = torch.empty(300, dtype=torch.int8, device="cuda")
    work = torch.empty(tk.workspace_size(20, 300), dtype=torch.uint8, device="cuda")
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            assert tk.zgeev_batched(A, 20, 300, eig, work, flag, handle, s) == 0
    s.synchronize()
    ref, _ = oracle.lapack_zgeev(M)
    assert oracle.max_relerr(ref, eig.cpu().numpy()) <= TOL_LAPACK
    # A is not clobbered
    np.testing.assert_array_equal(A.transpose(-1, -2).cpu().numpy(), M)


def test_deterministic_across_calls_batches_and_graph():
    # The same matrix gives the same bits however it is batched: alone, inside
    # a larger batch, in a later call, and through the ragged path (same
    # kernel code for n = 64) -- no atomics or reductions depend on scheduling.
    from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
    rng = np.random.default_rng(12)
    n = 64
    M = rand(rng, 500, n)
    h = tk.Handle()
    e1, _ = gpu_eigs(M, h)
    e2, _ = gpu_eigs(M, h)
    np.testing.assert_array_equal(e1, e2)
    e3, _ = gpu_eigs(M[123:124], h)
    np.testing.assert_array_equal(e1[123:124], e3)
    flat, offs = pack(list(M[:50]))
    rb = RaggedBatch(offs, handle=h)
    for graph in (False, True, True):
        eg, fg = rb.solve(torch.tensor(flat, device="cuda"), graph=graph)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(eg.cpu().numpy().reshape(50, n), e1[:50])
    h.close()


@pytest.mark.parametrize("n", [33, 40, 48, 49, 56, 64])
def test_register_hessenberg_paths(handle, n):
    """The rare paths of the register-resident Hessenberg reduction (n > 32)
    against LAPACK zgeev(x) on the same matrices:
    * a column whose tail has a sum of squares below 1e-290 while the matrix
      norm is in range (the rescaled second pass of the reflector's norm), at
      columns on both sides of the 16-column blocks;
    * block-diagonal matrices (columns with zero tails: tau = 0 between
      ordinary reflectors), with block boundaries on and off the 16-row
      blocks of the register grid;
    * upper Hessenberg input with complex sub-diagonal (zero tails, alpha not
      real: beta = -|alpha|, v = e_1).
    Bar: condition-scaled error <= 100 (units of kappa_i eps ||A||_F) and
    absolute error <= 1e-13 ||A||_F after matching."""
    rng = np.random.default_rng(90 + n)
    cases = []
    # upper Hessenberg but for column j's tail (rows j+2..), of size 1e-150:
    # the reflectors before step j only change phases (v = e_1), so step j
    # sees the tiny tail
    G = rand(rng, 6, n)
    for b, j in enumerate((0, 1, 15, 16, 17, n - 3)):
        tail = G[b, j + 2:, j].copy()
        G[b] = np.triu(G[b], -1)
        G[b, j + 2:, j] = tail * 1e-150
    cases.append(("tiny tail", G))
    B = rand(rng, 6, n)
    for b, cut in enumerate((1, 15, 16, 17, n // 2, n - 2)):
        B[b, cut:, :cut] = 0
        B[b, :cut, cut:] = 0
    cases.append(("block diagonal", B))
    Hh = rand(rng, 4, n)
    for b in range(4):
        Hh[b] = np.triu(Hh[b], -1)
    cases.append(("Hessenberg input", Hh))
    for name, M in cases:
        e, flag = gpu_eigs(M, handle)
        assert not flag.any(), name
        cs = oracle.condscaled_err(M, e)
        assert cs.max() <= 100, (name, cs.max())
        ref, _ = oracle.lapack_zgeev(M)
        for b in range(M.shape[0]):
            perm = oracle.match(ref[b], e[b])
            assert np.abs(e[b][perm] - ref[b]).max() <= 1e-13 * np.linalg.norm(M[b]), (name, b)
