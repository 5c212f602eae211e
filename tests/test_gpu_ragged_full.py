"""GPU parity of the ragged path at the full mixed-size configuration, where
every size class holds many times more work items than the GPU holds
resident, so the persistent kernels' wavefronts pull work from their queues
over many rounds.

* Eigenvalues: 200,000 matrices, n ~ U{4..64} (seed 1, the benchmark's
  batch).  Every matrix is compared bit for bit with the same-size entry
  called once per size bucket with the variants whose arithmetic the ragged
  kernels share (n > 32: the 256-thread Hessenberg and the two-bulge QR; n
  17..32 and n <= 16: the split lane-group kernels with the ragged
  configuration's lanes per matrix); every flag is 0; a 3,000-matrix sample
  is checked against host LAPACK zgeev (max rel err <= 1e-10 after optimal
  matching).
* Differentiation: 24,000 matrices, n ~ U{1..64}: the ragged Schur forward
  and backward against the same-size Schur forward and backward per bucket,
  bit for bit (eigenvalues, T, U, dL/dA), and a sample of dL/dA against the
  oracle's adjoint (<= 1e-9 relative to max |dL/dA|).
"""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.autograd import backward_batched, schur_batched
from tokura_zgeev_batched_amd.ragged import RaggedBatch

pytestmark = pytest.mark.gpu


def ragged_batch_on_device(sizes, seed):
    """Flat column-major ragged buffer on the device and, per n, the bucket of
    its matrices (indices, (count, n*n) column-major rows)."""
    B = len(sizes)
    offs = np.zeros(B + 1, np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)
    g = torch.Generator(device="cuda").manual_seed(seed)
    flat = torch.empty(int(offs[-1]), dtype=torch.complex128, device="cuda")
    offs_t = torch.tensor(offs[:-1], device="cuda")
    sizes_t = torch.tensor(sizes, device="cuda")
    buckets = {}
    for n in np.unique(sizes):
        n = int(n)
        idx = torch.nonzero(sizes_t == n).flatten()
        cnt = int(idx.numel())
        A = torch.complex(torch.randn(cnt, n * n, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(cnt, n * n, generator=g, device="cuda", dtype=torch.float64))
        pos = (offs_t[idx][:, None] + torch.arange(n * n, device="cuda")[None]).flatten()
        flat[pos] = A.flatten()
        buckets[n] = (idx, A)
    return flat, offs, buckets


def matched_plain_handle(lanes_b, lanes_c):
    """Same-size handle whose variants share the ragged kernels' arithmetic."""
    h = tk.Handle()
    for n in range(1, 65):
        if n > 32:
            v = 20                                   # 256-thread Hessenberg + two-bulge QR
        elif n > 16:
            v = 14 if lanes_b == 32 else 13          # split 32 x 32 / 32 x 16
        else:
            v = 27 if lanes_c == 8 else 26           # split 16 x 8 / 16 x 16
        h.set_variant(n, v)
    return h


@pytest.mark.parametrize("cfg", [None, 14])   # the handle's default; a wavefront cap on n 49..64
def test_ragged_full_config_every_queue_many_rounds(cfg):
    B = 200_000
    rng = np.random.default_rng(1)
    sizes = rng.integers(4, 65, B)
    flat, offs, buckets = ragged_batch_on_device(sizes, 1)
    h = tk.Handle()
    if cfg is not None:
        assert tk.lib().tokuraSetRaggedConfig(h.ptr, cfg) == 0
    rb = RaggedBatch(offs, handle=h)
    # resident matrices of the GPU at the largest class: well below each
    # class's count, so every queue serves many rounds
    per_class = np.bincount(np.digitize(sizes, [17, 33, 41, 49, 57, 61]))
    print(f"[case] ragged full config 200000 cfg={cfg}, class counts {per_class.tolist()}", flush=True)
    assert per_class.min() >= 6 * 1024
    eig, flag = rb.solve(flat, graph=True)
    torch.cuda.synchronize()
    assert int((flag != 0).sum()) == 0
    cfg, lanes = rb.config()
    hp = matched_plain_handle(lanes[2], lanes[3])
    eoff = torch.tensor(rb.eig_offsets[:-1], device="cuda")
    for n, (idx, A) in buckets.items():
        cnt = A.shape[0]
        e = torch.empty((cnt, n), dtype=torch.complex128, device="cuda")
        f = torch.empty(cnt, dtype=torch.int8, device="cuda")
        w = torch.empty(tk.workspace_size(n, cnt), dtype=torch.uint8, device="cuda")
        assert tk.zgeev_batched(A, n, cnt, e, w, f, hp) == 0
        got = eig[(eoff[idx][:, None] + torch.arange(n, device="cuda")[None])]
        assert torch.equal(got, e), n
        assert int((f != 0).sum()) == 0
    # host LAPACK on a sample
    pick = np.sort(rng.choice(B, size=3000, replace=False))
    e_host = eig.cpu().numpy()
    worst = 0.0
    for n in np.unique(sizes[pick]):
        sel = pick[sizes[pick] == n]
        M = np.stack([flat[offs[b]:offs[b + 1]].cpu().numpy().reshape(n, n).T for b in sel])
        ref, _ = oracle.lapack_zgeev(M)
        got = np.stack([e_host[rb.eig_offsets[b]:rb.eig_offsets[b + 1]] for b in sel])
        worst = max(worst, oracle.max_relerr(ref, got))
    print(f"ragged full config: 200000 bitwise vs per-bucket plain, sample max rel err vs LAPACK {worst:.2e}",
          flush=True)
    assert worst <= 1e-10


def test_ragged_schur_fwd_bwd_many_rounds():
    B = 24_000
    rng = np.random.default_rng(64)
    sizes = rng.integers(1, 65, B)
    flat, offs, buckets = ragged_batch_on_device(sizes, 64)
    rb = RaggedBatch(offs)
    print(f"[case] ragged schur fwd+bwd {B}", flush=True)
    eig, schur, flag = rb.schur(flat)
    g = torch.Generator(device="cuda").manual_seed(65)
    ne = int(rb.eig_offsets[-1])
    gflat = torch.complex(torch.randn(ne, generator=g, device="cuda", dtype=torch.float64),
                          torch.randn(ne, generator=g, device="cuda", dtype=torch.float64))
    dA = rb.backward(schur, gflat, flag)
    torch.cuda.synchronize()
    assert int((flag != 0).sum()) == 0
    eoff = torch.tensor(rb.eig_offsets[:-1], device="cuda")
    moff = torch.tensor(offs[:-1], device="cuda")
    for n, (idx, A) in buckets.items():
        cnt = A.shape[0]
        e1, s1, f1 = schur_batched(A, n, cnt)
        gi = gflat[(eoff[idx][:, None] + torch.arange(n, device="cuda")[None])]
        d1, _ = backward_batched(n, cnt, gi, s1, f1)
        torch.cuda.synchronize()
        assert torch.equal(eig[(eoff[idx][:, None] + torch.arange(n, device="cuda")[None])], e1), n
        spos = 2 * moff[idx][:, None] + torch.arange(2 * n * n, device="cuda")[None]
        assert torch.equal(schur[spos], s1.reshape(cnt, -1)), n
        dpos = moff[idx][:, None] + torch.arange(n * n, device="cuda")[None]
        assert torch.equal(dA[dpos], d1.reshape(cnt, -1)), n
    # a sample of gradients against the oracle
    pick = rng.choice(B, size=40, replace=False)
    e_h, g_h, d_h = eig.cpu().numpy(), gflat.cpu().numpy(), dA.cpu().numpy()
    for b in pick:
        n = int(sizes[b])
        o, eo = int(offs[b]), int(rb.eig_offsets[b])
        M = flat[o:o + n * n].cpu().numpy().reshape(n, n).T
        G, w = oracle.method_adjoint(M, np.zeros(n))
        perm = oracle.match(e_h[eo:eo + n], w)
        g_o = np.zeros(n, complex)
        g_o[perm] = g_h[eo:eo + n]
        Go, _ = oracle.method_adjoint(M, g_o)
        Gg = d_h[o:o + n * n].reshape(n, n).T
        assert np.abs(Gg - Go).max() <= 1e-9 * max(1.0, np.abs(Go).max()), (b, n)
