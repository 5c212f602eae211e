"""GPU parity of the ragged path (persistent work-queue kernel + hipGraph).

Each matrix of a mixed-size batch must match host LAPACK zgeev (max rel err
<= 1e-10 after optimal matching) and the same-size path on the same matrix
(<= 1e-12: same method, only the lane split of reductions differs).
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack

pytestmark = pytest.mark.gpu


def rand_mats(rng, sizes):
    return [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for n in sizes]


def check_against_lapack(mats, eig_flat, eoff, tol=1e-10):
    by_n = {}
    for b, m in enumerate(mats):
        by_n.setdefault(m.shape[0], []).append(b)
    worst = 0.0
    for n, idx in by_n.items():
        M = np.stack([mats[b] for b in idx])
        ref, _ = oracle.lapack_zgeev(M)
        got = np.stack([eig_flat[eoff[b]:eoff[b + 1]] for b in idx])
        worst = max(worst, oracle.max_relerr(ref, got))
    assert worst <= tol, worst
    return worst


@pytest.mark.parametrize("cfg", list(range(16)))
def test_ragged_every_config(cfg):
    rng = np.random.default_rng(100 + cfg)
    sizes = rng.integers(1, 65, 700)
    mats = rand_mats(rng, sizes)
    flat, offs = pack(mats)
    h = tk.Handle()
    assert tk.lib().tokuraSetRaggedConfig(h.ptr, cfg) == 0
    print(f"[case] ragged cfg={cfg} batch=700", flush=True)
    rb = RaggedBatch(offs, handle=h)
    A = torch.tensor(flat, device="cuda")
    eig, flag = rb.solve(A)
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0
    check_against_lapack(mats, eig.cpu().numpy(), rb.eig_offsets)
    h.close()


def test_ragged_matches_plain_path(handle):
    rng = np.random.default_rng(7)
    sizes = np.repeat(np.arange(1, 65), 5)
    rng.shuffle(sizes)
    mats = rand_mats(rng, sizes)
    flat, offs = pack(mats)
    rb = RaggedBatch(offs, handle=handle)
    print("[case] ragged vs plain", flush=True)
    eig, flag = rb.solve(torch.tensor(flat, device="cuda"))
    e = eig.cpu().numpy()
    for n in range(1, 65):
        idx = [b for b in range(len(mats)) if sizes[b] == n]
        M = np.stack([mats[b] for b in idx])
        pe, pf = tk.eigvals(torch.tensor(M, device="cuda"), handle)
        got = np.stack([e[rb.eig_offsets[b]:rb.eig_offsets[b + 1]] for b in idx])
        assert oracle.max_relerr(pe.cpu().numpy(), got) <= 1e-12, n


def test_ragged_graph_replays():
    rng = np.random.default_rng(8)
    sizes = rng.integers(4, 65, 400)
    mats = rand_mats(rng, sizes)
    flat, offs = pack(mats)
    h = tk.Handle()
    rb = RaggedBatch(offs, handle=h)
    A = torch.tensor(flat, device="cuda")
    eager, eflag = rb.solve(A)
    eager = eager.clone()
    torch.cuda.synchronize()
    eig = torch.empty_like(eager)
    flag = torch.empty_like(eflag)
    for rep in range(4):
        print(f"[case] graph replay {rep}", flush=True)
        eig.fill_(0)
        flag.fill_(7)
        rb.solve(A, eig, flag, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(eig, eager), rep
        assert int(flag.abs().sum()) == 0
        # eager work between replays
        x = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
    check_against_lapack(mats, eager.cpu().numpy(), rb.eig_offsets)
    h.close()


def test_ragged_flags_and_tiny():
    rng = np.random.default_rng(9)
    sizes = [1, 1, 2, 3, 64, 16, 17, 33, 1, 50]
    mats = rand_mats(rng, sizes)
    mats[5][3, 4] = np.nan
    offsets = np.zeros(len(mats) + 1, np.int64)
    flat2 = []
    cur = 0
    for b, m in enumerate(mats):
        offsets[b] = cur
        flat2.append(np.asarray(m).T.reshape(-1))
        cur += m.size
    offsets[len(mats)] = cur
    rb = RaggedBatch(offsets)
    A = torch.tensor(np.concatenate(flat2), device="cuda")
    eig, flag = rb.solve(A)
    f = flag.cpu().numpy()
    assert f[5] == 2 and np.isnan(eig.cpu().numpy()[rb.eig_offsets[5]:rb.eig_offsets[6]]).all()
    assert (np.delete(f, 5) == 0).all()
    e = eig.cpu().numpy()
    for b, m in enumerate(mats):
        if b == 5:
            continue
        ref, _ = oracle.lapack_zgeev(m[None])
        assert oracle.max_relerr(ref, e[None, rb.eig_offsets[b]:rb.eig_offsets[b + 1]]) <= 1e-10


def test_ragged_autotune_and_bad_offsets(handle):
    rng = np.random.default_rng(10)
    sizes = rng.integers(4, 65, 300)
    mats = rand_mats(rng, sizes)
    flat, offs = pack(mats)
    rb = RaggedBatch(offs, handle=handle)
    A = torch.tensor(flat, device="cuda")
    cfg, lanes = rb.autotune(A)
    assert 0 <= cfg < 13
    eig, flag = rb.solve(A)
    check_against_lapack(mats, eig.cpu().numpy(), rb.eig_offsets)
    bad = offs.copy()
    bad[2] += 1
    with pytest.raises(tk.TokuraError):
        RaggedBatch(bad, handle=handle)
    L = tk.lib()
    P = ctypes.c_void_p
    w = torch.empty(16, dtype=torch.uint8, device="cuda")
    # a work buffer that was never prepared
    assert L.tokura_zgeev_ragged_gpu(handle.ptr, 3, P(A.data_ptr()), P(w.data_ptr()), P(w.data_ptr()),
                                     P(w.data_ptr()), P(0)) == -8


def test_ragged_stale_schedule_flags_4():
    """A work buffer whose schedule was written over since prepare (as a freed
    buffer reallocated at the same address would be): no matrix is computed,
    every flag is 4, for the plain and the graph entry; a new prepare
    restores it."""
    rng = np.random.default_rng(11)
    sizes = rng.integers(1, 65, 300)
    mats = rand_mats(rng, sizes)
    flat, offs = pack(mats)
    h = tk.Handle()
    rb = RaggedBatch(offs, handle=h)
    A = torch.tensor(flat, device="cuda")
    e0, f0 = rb.solve(A, graph=True)
    e0 = e0.clone()
    torch.cuda.synchronize()
    assert int(f0.abs().sum()) == 0
    rb.work[:4096].zero_()
    for graph in (False, True):
        eig, flag = rb.solve(A, graph=graph)
        torch.cuda.synchronize()
        assert (flag.cpu().numpy() == 4).all(), graph
    rb.prepare()
    eig, flag = rb.solve(A, graph=True)
    torch.cuda.synchronize()
    assert int(flag.abs().sum()) == 0 and torch.equal(eig, e0)
    h.close()


def test_ragged_reprepare_other_offsets_same_buffer():
    """Preparing the same work buffer again for other offsets (same batch)
    replaces the schedule; the cached graph is captured again."""
    rng = np.random.default_rng(13)
    h = tk.Handle()
    sizes1 = rng.integers(1, 65, 200)
    mats1 = rand_mats(rng, sizes1)
    flat1, offs1 = pack(mats1)
    rb = RaggedBatch(offs1, handle=h)
    e1, f1 = rb.solve(torch.tensor(flat1, device="cuda"), graph=True)
    check_against_lapack(mats1, e1.cpu().numpy(), rb.eig_offsets)
    sizes2 = np.sort(sizes1)            # same batch, other sizes per position
    mats2 = rand_mats(rng, sizes2)
    flat2, offs2 = pack(mats2)
    rb2 = RaggedBatch.__new__(RaggedBatch)
    rb2.__dict__.update(rb.__dict__)
    rb2.offsets = offs2
    rb2.sizes = sizes2
    rb2.eig_offsets = np.concatenate([[0], np.cumsum(sizes2)]).astype(np.int64)
    ws = int(tk.lib().tokura_get_zgeev_ragged_worksize(rb2.batch, offs2.ctypes.data_as(ctypes.c_void_p)))
    if ws > rb.work.numel():
        rb2.work = torch.empty(ws, dtype=torch.uint8, device="cuda")
    rb2.prepare()
    e2, f2 = rb2.solve(torch.tensor(flat2, device="cuda"), graph=True)
    torch.cuda.synchronize()
    assert int(f2.abs().sum()) == 0
    check_against_lapack(mats2, e2.cpu().numpy(), rb2.eig_offsets)
    h.close()


def test_ragged_wrapper_rejects_bad_buffers():
    """Caller-supplied output and cotangent buffers are checked before any
    launch: too short, wrong dtype, on the host or on another device."""
    rng = np.random.default_rng(7)
    mats = rand_mats(rng, [5, 12, 40])
    flat, offs = pack(mats)
    A = torch.from_numpy(flat).cuda()
    rb = RaggedBatch(offs, handle=tk.Handle())
    ne = int(rb.eig_offsets[-1])
    bad = [dict(eig=torch.empty(ne - 1, dtype=torch.complex128, device="cuda")),
           dict(eig=torch.empty(ne, dtype=torch.complex64, device="cuda")),
           dict(eig=torch.empty(ne, dtype=torch.complex128)),
           dict(flag=torch.empty(rb.batch, dtype=torch.int32, device="cuda")),
           dict(flag=torch.empty(rb.batch - 1, dtype=torch.int8, device="cuda"))]
    for kw in bad:
        with pytest.raises(tk.TokuraError):
            rb.solve(A, **kw)
    eig, schur, flag = rb.schur(A)
    with pytest.raises(tk.TokuraError):
        rb.backward(schur, torch.ones(ne, dtype=torch.complex128), flag)          # host gbar
    with pytest.raises(tk.TokuraError):
        rb.backward(schur, torch.ones(ne - 1, dtype=torch.complex128, device="cuda"), flag)
    with pytest.raises(tk.TokuraError):
        rb.backward(schur[:-1], torch.ones(ne, dtype=torch.complex128, device="cuda"), flag)
    dA = rb.backward(schur, torch.ones(ne, dtype=torch.complex128, device="cuda"), flag)
    assert dA.shape == A.shape and bool(torch.isfinite(dA).all())

