"""Benchmark of the batched eigenvalue path on one MI355X.

Default workload (the headline metric's configuration): 100,000 complex128
matrices, n = 64, entries Re/Im iid N(0,1) (torch generator on the device,
seed 0), column-major contiguous, eigenvalues only through the C ABI call
tokura_zgeev_batched_gpu.  A step is one call on the whole batch, inputs
resident in HBM.  Prints ONE JSON line.

Roofline: FLOP model F(n) = 40 n^3 per eigenvalues-only solve (Golub & Van
Loan ~10 n^3 real for Hessenberg + QR, x4 for complex); the Hessenberg share
is 40/3 n^3 (LAWN 41 ZGEHRD), so the QR kernel is charged 80/3 n^3.  Kernel
durations come from HIP events the library records on the launch stream
(tokuraSetStageTiming).  Peak: MI355X FP64 vector 78.6 TFLOP/s.

CPU baseline: host LAPACK zgeev (OpenBLAS, one thread per matrix, OpenMP over
matrices: the reference harness's scheme) on a time-bounded sample of the
same matrices; the sample's GPU eigenvalues are checked against it (max rel
error after optimal matching).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import tokura_zgeev_batched_amd as tk  # noqa: E402

PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# Per-matrix figures of the n = 64 QR kernels from rocprofv3 counter passes
# over one 20,000-matrix call (FETCH_SIZE / WRITE_SIZE each scaled by its
# reading for a 1 GiB device copy: FETCH x2.0, WRITE x1.0; SQ_INSTS_VALU_
# {FMA,MUL,ADD,TRANS}_F64 x 64 lanes per wave instruction, an FMA counted as 2):
#  - chunk variants 37-39 (k_qr_chunk<8>): 34,573 B read + 1,071 B written;
#    64 * (2 * 301,516 + 153,109 + 69,861 + 13,011) FP64 lane operations.
# The pool variants 29-36 and the fixed variant 20 run the same two-bulge
# step and move the same bytes (34,563 B read + 1,086 B written); their FP64
# counts were taken before the step's current form, so they report none.
# The FP64 count includes the serial reflector chain that every lane computes
# redundantly, ~7.7x the 80/3 n^3 model.
QR64_COUNTS = {v: (35_644.0, 5.3697e7) for v in (37, 38, 39)}


def flops_per_solve(n: int) -> float:
    return 40.0 * n ** 3


def qr_flops_per_solve(n: int) -> float:
    return 80.0 / 3.0 * n ** 3


def bytes_per_solve(n: int) -> float:
    return 16.0 * n * n + 16.0 * n + 1.0


def make_batch(B: int, n: int, seed: int, dev) -> torch.Tensor:
    """(B, n*n) complex128, each row one matrix in column-major order."""
    g = torch.Generator(device=dev).manual_seed(seed)
    re = torch.randn(B, n * n, generator=g, device=dev, dtype=torch.float64)
    im = torch.randn(B, n * n, generator=g, device=dev, dtype=torch.float64)
    A = torch.complex(re, im)
    del re, im
    return A


def cpu_baseline(A_host: np.ndarray, n: int, budget_s: float):
    """Times LAPACK zgeev on chunks of the sample until budget_s elapses."""
    import oracle
    M = A_host.reshape(-1, n, n).transpose(0, 2, 1)   # column-major rows -> M[b, r, c]
    done, t0 = 0, time.perf_counter()
    eig_chunks = []
    chunk = 1024
    while done < M.shape[0] and time.perf_counter() - t0 < budget_s:
        e, info = oracle.lapack_zgeev(M[done:done + chunk])
        eig_chunks.append(e)
        done += min(chunk, M.shape[0] - done)
    dt = time.perf_counter() - t0
    return done / dt, done, np.concatenate(eig_chunks), oracle.max_threads()


def timed(fn, steps: int, warmup: int) -> float:
    """Seconds per call of fn, HIP events on the current stream."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / steps


def bench_ragged(dev, B=200_000, lo=4, hi=64, steps=5, seed=1):
    """Mixed sizes n ~ U{lo..hi}: persistent work-queue kernel (hipGraph) vs
    the same-size path called once per size bucket on the same matrices."""
    from tokura_zgeev_batched_amd.ragged import RaggedBatch
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi + 1, B)
    offs = np.zeros(B + 1, np.int64)
    offs[1:] = np.cumsum(sizes.astype(np.int64) ** 2)
    g = torch.Generator(device=dev).manual_seed(seed)
    flat = torch.empty(int(offs[-1]), dtype=torch.complex128, device=dev)
    buckets = {}
    offs_t = torch.tensor(offs[:-1], device=dev)
    sizes_t = torch.tensor(sizes, device=dev)
    for n in range(lo, hi + 1):
        idx = torch.nonzero(sizes_t == n).flatten()
        if idx.numel() == 0:
            continue
        cnt = int(idx.numel())
        A = torch.complex(torch.randn(cnt, n * n, generator=g, device=dev, dtype=torch.float64),
                          torch.randn(cnt, n * n, generator=g, device=dev, dtype=torch.float64))
        pos = (offs_t[idx][:, None] + torch.arange(n * n, device=dev)[None]).flatten()
        flat[pos] = A.flatten()
        buckets[n] = (A, torch.empty((cnt, n), dtype=torch.complex128, device=dev),
                      torch.empty(cnt, dtype=torch.int8, device=dev),
                      torch.empty(tk.workspace_size(n, cnt), dtype=torch.uint8, device=dev))
    h = tk.Handle()
    h.autotune(lo, hi, 4096)            # the per-bucket baseline gets tuned variants too
    rb = RaggedBatch(offs, handle=h)
    cfg, lanes = rb.autotune(flat)
    eig = torch.empty(int(rb.eig_offsets[-1]), dtype=torch.complex128, device=dev)
    flag = torch.empty(B, dtype=torch.int8, device=dev)
    t_rag = timed(lambda: rb.solve(flat, eig, flag, graph=True), steps, 1)
    nbad = int((flag != 0).sum())

    def plain_buckets():
        for n, (A, e, f, w) in buckets.items():
            rc = tk.zgeev_batched(A, n, A.shape[0], e, w, f, h)
            if rc != 0:
                raise RuntimeError(f"plain call n={n} returned {rc}")
    t_plain = timed(plain_buckets, steps, 1)
    # the two paths agree on every matrix of one bucket
    n0 = hi
    idx = np.nonzero(sizes == n0)[0]
    got = np.stack([eig[rb.eig_offsets[b]:rb.eig_offsets[b + 1]].cpu().numpy() for b in idx[:200]])
    ref = buckets[n0][1][:200].cpu().numpy()
    import oracle
    agree = oracle.max_relerr(ref, got)
    flops = float(np.sum(40.0 * sizes.astype(np.float64) ** 3))
    return {"workload": f"ragged, batch={B}, n~U{{{lo}..{hi}}} seed {seed}, int64 offsets",
            "ragged_matrices_per_s": B / t_rag, "ragged_ms": t_rag * 1e3,
            "plain_per_bucket_matrices_per_s": B / t_plain, "plain_per_bucket_ms": t_plain * 1e3,
            "speedup_vs_plain_per_bucket": t_plain / t_rag,
            "ragged_config": {"id": cfg, "lanes_per_matrix_A1_A2_B_C": lanes, "hess_grid_multiplier": rb.hess_grid()},
            "model_tflops": flops / t_rag / 1e12, "flags_nonzero": nbad,
            "max_rel_diff_ragged_vs_plain_n64_sample": agree}


def bench_hard(dev, B=100_000, lo=8, hi=64, steps=3, seed=5, sample=1200):
    """Hard spectra (half companions, half clustered with kappa(V) ~ 1e4),
    ragged n ~ U{lo..hi}, through the persistent ragged path (hipGraph)."""
    from tokura_zgeev_batched_amd.ragged import RaggedBatch, pack
    from tokura_zgeev_batched_amd.workloads import hard_batch
    import oracle
    mats, kinds = hard_batch(B, lo, hi, seed)
    flat, offs = pack(mats)
    h = tk.Handle()
    rb = RaggedBatch(offs, handle=h)
    A = torch.tensor(flat, device=dev)
    rb.autotune(A)
    eig = torch.empty(int(rb.eig_offsets[-1]), dtype=torch.complex128, device=dev)
    flag = torch.empty(B, dtype=torch.int8, device=dev)
    t = timed(lambda: rb.solve(A, eig, flag, graph=True), steps, 1)
    f = flag.cpu().numpy().astype(np.int64)
    hist = {str(k): int(v) for k, v in enumerate(np.bincount(f, minlength=3))}
    e = eig.cpu().numpy()
    rng = np.random.default_rng(seed + 1)
    pick = np.sort(rng.choice(B, size=min(sample, B), replace=False))
    sizes = np.array([mats[b].shape[0] for b in pick])
    worst = {"companion": 0.0, "clustered": 0.0}
    for n in np.unique(sizes):
        idx = pick[sizes == n]
        M = np.stack([mats[b] for b in idx])
        got = np.stack([e[rb.eig_offsets[b]:rb.eig_offsets[b + 1]] for b in idx])
        err = oracle.condscaled_err(M, got).max(axis=1)
        for b, x in zip(idx, err):
            worst[kinds[b]] = max(worst[kinds[b]], float(x))
    flops = float(sum(40.0 * m.shape[0] ** 3 for m in mats))
    # ragged fwd (Schur vectors) + bwd (adjoint) on the same batch, cotangents iid complex N(0,1)
    g = torch.Generator(device=dev).manual_seed(seed + 2)
    ne = int(rb.eig_offsets[-1])
    gbar = torch.complex(torch.randn(ne, generator=g, device=dev, dtype=torch.float64),
                         torch.randn(ne, generator=g, device=dev, dtype=torch.float64))
    st = {}

    def fwd_schur():
        st["r"] = rb.schur(A)

    t_schur = timed(fwd_schur, steps, 1)
    _, schur, fl = st["r"]
    t_bwd = timed(lambda: rb.backward(schur, gbar, fl), steps, 1)
    diff = {"fwd_schur_ms": t_schur * 1e3, "bwd_ms": t_bwd * 1e3, "fwd_bwd_solves_per_s": B / (t_schur + t_bwd),
            "fwd_bwd_over_fwd_eigvals": (t_schur + t_bwd) / t, "flags_nonzero": int((fl != 0).sum())}
    del schur, st
    return {"workload": f"hard spectra, batch={B}, n~U{{{lo}..{hi}}} seed {seed}: half companions of monic "
                        "polynomials with iid complex N(0,1) coefficients, half V diag(lambda) V^-1 with 3 "
                        "clusters of spread 1e-3 and kappa(V)=1e4",
            "matrices_per_s": B / t, "ms": t * 1e3, "model_tflops": flops / t / 1e12, "flag_histogram": hist,
            "condition_scaled_err_max": worst,
            "condition_scaled_err_def": "|lambda - lambda_zgeevx| * rcond / (eps ||A||_F), unbalanced zgeevx",
            "accuracy_sample": int(len(pick)), "ragged_fwd_bwd": diff}


def bench_end_to_end(dev, sizes=(62, 64), B=16_000, chunk=8192, nstreams=8, seed=7):
    """Host-resident batches through the streaming pipeline (H2D + solve + D2H
    on 8 streams, 8192-matrix chunks), the reference harness's measurement
    mode (stream_test_with_MKL/main.cu:245-296): PCIe-inclusive matrices/s."""
    from tokura_zgeev_batched_amd.pipeline import pinned_colmajor, stream_eigvals
    out = {}
    g = torch.Generator().manual_seed(seed)
    for n in sizes:
        A = pinned_colmajor(B, n)
        A.copy_(torch.complex(torch.randn(B, n * n, generator=g, dtype=torch.float64),
                              torch.randn(B, n * n, generator=g, dtype=torch.float64)))
        stream_eigvals(A, n, chunk, nstreams)   # warm-up
        t0 = time.perf_counter()
        eig, flag = stream_eigvals(A, n, chunk, nstreams)
        dt = time.perf_counter() - t0
        out[f"n{n}"] = {"matrices_per_s_incl_pcie": B / dt, "seconds_for_16000": dt,
                        "flags_nonzero": int((flag != 0).sum())}
    out["workload"] = (f"host-resident batch of {B} matrices per n, {nstreams} streams x {chunk}-matrix chunks, "
                       "H2D + solve + D2H (PCIe-inclusive)")
    out["reference_published_n62_matrices_per_s_incl_transfers"] = 3225.0
    return out


def bench_diff(dev, B=50_000, n=32, steps=3, seed=3):
    """Forward (eigenvalues only) vs forward with Schur vectors + adjoint, and
    the recompute mode: the same eigenvalue forward, then a backward that
    rebuilds the Schur form in the workspace and pairs its eigenvalues with
    the forward's."""
    from tokura_zgeev_batched_amd.autograd import backward_batched, backward_recompute, schur_batched
    g = torch.Generator(device=dev).manual_seed(seed)
    A = make_batch(B, n, seed, dev)
    gbar = torch.complex(torch.randn(B, n, generator=g, device=dev, dtype=torch.float64),
                         torch.randn(B, n, generator=g, device=dev, dtype=torch.float64))
    h = tk.Handle()
    eig = torch.empty((B, n), dtype=torch.complex128, device=dev)
    flag = torch.empty(B, dtype=torch.int8, device=dev)
    work = torch.empty(tk.workspace_size(n, B), dtype=torch.uint8, device=dev)
    t_fwd = timed(lambda: tk.zgeev_batched(A, n, B, eig, work, flag, h), steps, 1)
    state = {}

    def fwd_schur():
        state["r"] = schur_batched(A, n, B, h)

    t_schur = timed(fwd_schur, steps, 1)
    e_s, schur, fl = state["r"]

    def bwd():
        backward_batched(n, B, gbar, schur, fl, handle=h)

    t_bwd = timed(bwd, steps, 1)
    dA_store, _ = backward_batched(n, B, gbar, schur, fl, handle=h)
    del schur
    state.clear()
    torch.cuda.empty_cache()

    def bwd_recompute():
        state["rc"] = backward_recompute(n, B, A, eig, gbar, handle=h)

    t_bwd_rc = timed(bwd_recompute, steps, 1)
    _, fl_rc = state["rc"]
    # the recompute backward's Schur form comes from snapped shifts, so its
    # gradient agrees with the stored-Schur one to rounding: the largest
    # entry-wise difference relative to each matrix's largest gradient entry
    dA_rc, _ = backward_recompute(n, B, A, e_s, gbar, handle=h)
    scale = dA_store.abs().amax(dim=(1, 2)).clamp(min=1.0)
    rel = float(((dA_rc - dA_store).abs().amax(dim=(1, 2)) / scale).max())
    return {"workload": f"differentiation, batch={B}, n={n}, loss Re sum conj(g) lambda",
            "fwd_eigvals_ms": t_fwd * 1e3, "fwd_schur_ms": t_schur * 1e3, "bwd_ms": t_bwd * 1e3,
            "bwd_recompute_ms": t_bwd_rc * 1e3,
            "fwd_bwd_solves_per_s": B / (t_schur + t_bwd),
            "fwd_bwd_over_fwd_eigvals": (t_schur + t_bwd) / t_fwd,
            "fwd_bwd_over_fwd_schur": (t_schur + t_bwd) / t_schur,
            "recompute_fwd_bwd_solves_per_s": B / (t_fwd + t_bwd_rc),
            "recompute_fwd_bwd_over_fwd_eigvals": (t_fwd + t_bwd_rc) / t_fwd,
            "recompute_flags_nonzero": int((fl_rc != 0).sum()),
            "recompute_grad_max_rel_diff_vs_store": rel}


def bench_vectors(dev, cases=((32, 50_000), (64, 20_000)), steps=3, seed=11):
    """Eigenvalues + right and left eigenvectors (ZGEEV 'V','V' analogue)."""
    out = {"workload": "eigenvectors (left and right), same-size batches"}
    h = tk.Handle()
    for n, B in cases:
        A = make_batch(B, n, seed, dev)
        eig = torch.empty((B, n), dtype=torch.complex128, device=dev)
        flag = torch.empty(B, dtype=torch.int8, device=dev)
        VL = torch.empty((B, n, n), dtype=torch.complex128, device=dev)
        VR = torch.empty((B, n, n), dtype=torch.complex128, device=dev)

        def both():
            rc = tk.zgeev_vectors(A, n, B, eig, VL, VR, flag, h)
            if rc != 0:
                raise RuntimeError(f"tokura_zgeev_batched_vectors_gpu returned {rc}")

        work = torch.empty(tk.vectors_workspace_size(n, B), dtype=torch.uint8, device=dev)

        def right():          # T parked in the optional workspace
            rc = tk.zgeev_vectors(A, n, B, eig, None, VR, flag, h, work=work)
            if rc != 0:
                raise RuntimeError(f"tokura_zgeev_batched_vectors_gpu returned {rc}")

        def right_no_ws():    # one-kernel form, no workspace
            tk.zgeev_vectors(A, n, B, eig, None, VR, flag, h)

        t_both = timed(both, steps, 1)
        t_right = timed(right, steps, 1)
        t_right1 = timed(right_no_ws, steps, 1)
        out[f"n{n}"] = {"batch": B, "left_right_ms": t_both * 1e3, "right_only_ms": t_right * 1e3,
                        "right_only_no_workspace_ms": t_right1 * 1e3,
                        "matrices_per_s_left_right": B / t_both, "matrices_per_s_right": B / t_right,
                        "flags_nonzero": int((flag != 0).sum())}
        del A, eig, flag, VL, VR, work
        torch.cuda.empty_cache()
    return out


def bench_real(dev, cases=((64, 100_000), (32, 100_000), (16, 100_000)), steps=3, seed=19, cpu_seconds=4.0):
    """Real matrices (tokura_dgeev_batched_gpu, the DGEEV analogue): N(0,1)
    entries, eigenvalues only; accuracy vs host LAPACK dgeev on a sample and
    a bounded host LAPACK dgeev baseline (16 threads) at each n."""
    import oracle
    out = {"workload": "real matrices, same-size batches, eigenvalues only (DGEEV analogue)"}
    h = tk.Handle()
    for n, B in cases:
        g = torch.Generator(device=dev).manual_seed(seed)
        M = torch.randn(B, n, n, generator=g, device=dev, dtype=torch.float64)
        A = M.transpose(-1, -2).contiguous()
        wr = torch.empty((B, n), dtype=torch.float64, device=dev)
        wi = torch.empty_like(wr)
        flag = torch.empty(B, dtype=torch.int8, device=dev)
        work = torch.empty(tk.dgeev_workspace_size(n, B), dtype=torch.uint8, device=dev)

        def step():
            rc = tk.dgeev_batched(A, n, B, wr, wi, work, flag, h)
            if rc != 0:
                raise RuntimeError(f"tokura_dgeev_batched_gpu returned {rc}")

        t = timed(step, steps, 1)
        k = min(B, 2000)
        Ms = M[:k].cpu().numpy()
        w_gpu = (wr[:k] + 1j * wi[:k]).cpu().numpy()
        oracle.lapack_dgeev(Ms[:64])   # warm-up: library and thread-pool start-up stay out of the timing
        t0 = time.perf_counter()
        done = 0
        wl_parts = []
        while done < k and time.perf_counter() - t0 < cpu_seconds:
            wl, _ = oracle.lapack_dgeev(Ms[done:done + 500])
            wl_parts.append(wl)
            done += wl.shape[0]
        cpu_rate = done / (time.perf_counter() - t0)
        wl_all = np.concatenate(wl_parts)
        err = oracle.max_relerr(wl_all, w_gpu[:done])
        # real Gaussian matrices have real eigenvalues near 0, where any
        # backward-stable method's relative error grows: also in units of
        # kappa_i eps ||A||_F and relative to ||A||_F
        ks = min(done, 300)
        cs = float(oracle.condscaled_err(Ms[:ks], w_gpu[:ks]).max())
        na = max(float(np.abs(w_gpu[b][oracle.match(wl_all[b], w_gpu[b])] - wl_all[b]).max()
                       / np.linalg.norm(Ms[b])) for b in range(done))
        out[f"n{n}_b{B}"] = {"matrices_per_s": B / t, "ms": t * 1e3, "flags_nonzero": int((flag != 0).sum()),
                             # the bars of tests/test_gpu_real.py: condition-scaled error <= 100 and
                             # |lambda - lambda_lapack| <= 1e-13 ||A||_F; the plain relative error is
                             # inflated by real eigenvalues near 0 (any backward-stable method's is)
                             "accuracy_metric": "condition_scaled_err_max <= 100 (units of kappa_i eps ||A||_F) "
                                                "and max_abs_err_over_norm_A <= 1e-13",
                             "condition_scaled_err_max": cs, "max_abs_err_over_norm_A": na,
                             "rel_err_vs_lapack_dgeev_unscaled": err, "accuracy_sample": done,
                             "cpu_lapack_dgeev_matrices_per_s": cpu_rate, "cpu_threads": oracle.max_threads(),
                             "model_tflops_10n3": 10.0 * n ** 3 * B / t / 1e12}
        del M, A, wr, wi, flag, work
        torch.cuda.empty_cache()
    return out


def bench_small(dev, cases=((16, 10_000), (16, 200_000), (32, 100_000)), steps=5, seed=13):
    """Plain same-size path at small n (the reference's CPU-runnable case is
    n = 16 x 10,000), autotuned per n, accuracy vs host LAPACK zgeev."""
    import oracle
    out = {"workload": "plain same-size zgeev at small n, eigenvalues only"}
    h = tk.Handle()
    for n, B in cases:
        A = make_batch(B, n, seed, dev)
        eig = torch.empty((B, n), dtype=torch.complex128, device=dev)
        flag = torch.empty(B, dtype=torch.int8, device=dev)
        work = torch.empty(tk.workspace_size(n, B), dtype=torch.uint8, device=dev)
        h.autotune(n, n, min(B, 50_000))
        t = timed(lambda: tk.zgeev_batched(A, n, B, eig, work, flag, h), steps, 1)
        k = min(B, 4000)
        err = oracle.max_relerr(oracle.lapack_zgeev(A[:k].reshape(k, n, n).transpose(-1, -2).cpu().numpy())[0],
                                eig[:k].cpu().numpy())
        out[f"n{n}_b{B}"] = {"matrices_per_s": B / t, "ms": t * 1e3, "variant": h.launch_config(n)[0],
                             "max_rel_err_vs_lapack": err, "flags_nonzero": int((flag != 0).sum())}
        del A, eig, flag, work
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the ragged and differentiation legs")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch table")
    ap.add_argument("--only", choices=["ragged", "diff", "hard", "e2e", "vectors", "small", "real"], default=None,
                    help="run only this extra leg (for profiling); prints its JSON object")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the batched eigenvalue path shards by matrices "
              "with no exchange, so a multi-GPU figure is N x the one-GPU value (not measured here).",
              file=sys.stderr)
        sys.exit(2)

    dev = torch.device("cuda:0")
    if args.only:
        fn = {"ragged": bench_ragged, "diff": bench_diff, "hard": bench_hard, "e2e": bench_end_to_end,
              "vectors": bench_vectors, "small": bench_small, "real": bench_real}[args.only]
        print(json.dumps({args.only: fn(dev)}), flush=True)
        return
    n, B = args.n, args.batch
    A = make_batch(B, n, 0, dev)
    eig = torch.empty((B, n), dtype=torch.complex128, device=dev)
    flag = torch.empty((B,), dtype=torch.int8, device=dev)
    work = torch.empty((tk.workspace_size(n, B),), dtype=torch.uint8, device=dev)
    h = tk.Handle()
    if not args.no_tune:
        h.autotune(n, n, min(B, 32768))   # runtime autotuner picks the kernel variant for n
    variant, lpm, mpb = h.launch_config(n)
    stream = torch.cuda.current_stream()

    def step():
        rc = tk.zgeev_batched(A, n, B, eig, work, flag, h, stream)
        if rc != 0:
            raise RuntimeError(f"tokura_zgeev_batched_gpu returned {rc}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    h.set_stage_timing(True)
    h.stage_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    calls, hess_ms, qr_ms = h.stage_times()
    h.set_stage_timing(False)
    if args.copy_probe_mib > 0:
        src = torch.empty(args.copy_probe_mib * (1 << 20), dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    nflag = int((flag != 0).sum())
    value = B * args.steps / wall
    ms_step = wall * 1e3 / args.steps
    avg_qr_s = qr_ms / calls / 1e3
    avg_hess_s = hess_ms / calls / 1e3
    qr_achieved = qr_flops_per_solve(n) * B / avg_qr_s
    path_achieved = flops_per_solve(n) * B / ((qr_ms + hess_ms) / calls / 1e3)

    result = {
        "metric": "eigensolves/s at n=64 (fp64, err<=1e-10 vs LAPACK zgeev)",
        "value": value,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: complex128 Re/Im iid N(0,1), torch cuda generator seed 0, column-major",
        "config": {"workload": f"plain same-size zgeev, n={n}, batch={B}, eigenvalues only",
                   "n": n, "batch": B, "variant": variant, "lanes_per_matrix": lpm,
                   "matrices_per_workgroup": mpb},
        "roofline": {"bound": "fp64_vector", "achieved": qr_achieved / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": qr_achieved / PEAK_FP64,
                     "traffic": (QR64_COUNTS[variant][0] * B if n == 64 and variant in QR64_COUNTS else None),
                     "traffic_unit": "bytes per launch (rocprof FETCH_SIZE+WRITE_SIZE, calibrated)",
                     "algorithmic_bytes": (8.0 * (n - 1) * (n + 2) + 16.0 * n + 16.0 * n + 1.0) * B,
                     "kernel": ("QR stage (" + ("k_qr_chunk, one workgroup per CU, compacting LDS arena"
                                                if variant >= 37 else
                                                ("k_qr_pool, one workgroup per CU" if variant >= 29 else "k_qr_packed"))
                                + ") charged 80/3 n^3 flop/matrix"),
                     "compute_unit": "fp64 VALU (no MFMA is issued); MI355X fp64 vector peak 78.6 TFLOP/s",
                     "kernel_avg_ms": avg_qr_s * 1e3, "hessenberg_avg_ms": avg_hess_s * 1e3,
                     "path_achieved_tflops": path_achieved / 1e12,
                     "path_frac": path_achieved / PEAK_FP64,
                     # issued FP64 lane operations (rocprof counters per matrix, live kernel time)
                     "executed_fp64_tflops": (QR64_COUNTS[variant][1] * B / avg_qr_s / 1e12
                                              if n == 64 and variant in QR64_COUNTS else None),
                     "executed_fp64_frac": (QR64_COUNTS[variant][1] * B / avg_qr_s / PEAK_FP64
                                            if n == 64 and variant in QR64_COUNTS else None)},
        "flags_nonzero": nflag,
    }

    if not args.no_cpu:
        # accuracy of the timed output + CPU baseline on the same matrices
        sample = min(B, 60_000)
        A_host = A[:sample].cpu().numpy()
        cps, done, e_cpu, cores = cpu_baseline(A_host, n, args.cpu_seconds)
        import oracle
        err = oracle.max_relerr(e_cpu, eig[:done].cpu().numpy())
        result["cpu_baseline"] = {"value": cps, "unit": "matrices/s", "cores": cores, "kind": "port",
                                  "sample": f"first {done} matrices of the workload, host LAPACK zgeev "
                                            f"(OpenBLAS, 1 thread/matrix, OpenMP over {cores} threads)"}
        result["max_rel_err_vs_lapack"] = err
        result["accuracy_sample"] = done
    if not args.no_extra:
        del A, eig, flag, work
        torch.cuda.empty_cache()
        result["extra"] = {"ragged": bench_ragged(dev), "differentiation": bench_diff(dev),
                           "hard_spectra": bench_hard(dev), "end_to_end": bench_end_to_end(dev),
                           "eigenvectors": bench_vectors(dev), "small_n": bench_small(dev),
                           "real_matrices": bench_real(dev)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
