"""Hard spectra (companion matrices and clustered, ill-conditioned similarity
transforms, n ~ U{8..64}) on the GPU, checked with condition-scaled errors:
error_i * rcond_i / (eps ||A||_F) <= 100 against LAPACK zgeevx on the
unbalanced matrices (the GPU method does not balance, README.md:79)."""
import numpy as np
import pytest
import torch

import oracle
import tokura_zgeev_batched_amd as tk
from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
from tokura_zgeev_batched_amd.workloads import hard_batch

pytestmark = pytest.mark.gpu

BOUND = 100.0


def check_by_size(mats, eig_of):
    worst = {"companion": 0.0, "clustered": 0.0}
    sizes = np.array([m.shape[0] for m in mats])
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        M = np.stack([mats[b] for b in idx])
        got = np.stack([eig_of(b) for b in idx])
        err = oracle.condscaled_err(M, got).max(axis=1)
        for b, e in zip(idx, err):
            k = "companion" if b % 2 == 0 else "clustered"
            worst[k] = max(worst[k], float(e))
    print("condition-scaled worst:", worst, flush=True)
    assert max(worst.values()) <= BOUND
    return worst


def test_hard_spectra_ragged_graph():
    mats, kinds = hard_batch(1500, 8, 64, 21)
    flat, offs = pack(mats)
    rb = RaggedBatch(offs)
    print("[case] hard spectra ragged 1500", flush=True)
    eig, flag = rb.solve(torch.tensor(flat, device="cuda"), graph=True)
    torch.cuda.synchronize()
    f = flag.cpu().numpy()
    assert (f == 0).all(), np.bincount(f.astype(np.int64) & 0xFF)
    e = eig.cpu().numpy()
    check_by_size(mats, lambda b: e[rb.eig_offsets[b]:rb.eig_offsets[b + 1]])


def test_hard_spectra_plain_path(handle):
    mats, kinds = hard_batch(400, 40, 64, 22)
    sizes = np.array([m.shape[0] for m in mats])
    out = {}
    for n in np.unique(sizes):
        idx = np.nonzero(sizes == n)[0]
        print(f"[case] hard plain n={n}", flush=True)
        ev, fl = tk.eigvals(torch.tensor(np.stack([mats[b] for b in idx]), device="cuda"), handle)
        assert int(fl.abs().sum()) == 0
        for j, b in enumerate(idx):
            out[b] = ev[j].cpu().numpy()
    check_by_size(mats, lambda b: out[b])
