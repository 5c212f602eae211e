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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--batch", type=int, default=100_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the batched eigenvalue path shards by matrices "
              "with no exchange, so a multi-GPU figure is N x the one-GPU value (not measured here).",
              file=sys.stderr)
        sys.exit(2)

    dev = torch.device("cuda:0")
    n, B = args.n, args.batch
    A = make_batch(B, n, 0, dev)
    eig = torch.empty((B, n), dtype=torch.complex128, device=dev)
    flag = torch.empty((B,), dtype=torch.int8, device=dev)
    work = torch.empty((tk.workspace_size(n, B),), dtype=torch.uint8, device=dev)
    h = tk.Handle()
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
        "roofline": {"bound": "mfma", "achieved": qr_achieved / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": qr_achieved / PEAK_FP64,
                     "traffic": None,
                     "kernel": "QR stage (k_qr_packed) charged 80/3 n^3 flop/matrix",
                     "compute_unit": "fp64 VALU; MI355X fp64 vector and matrix peaks coincide (78.6 TFLOP/s)",
                     "kernel_avg_ms": avg_qr_s * 1e3, "hessenberg_avg_ms": avg_hess_s * 1e3,
                     "path_achieved_tflops": path_achieved / 1e12,
                     "path_frac": path_achieved / PEAK_FP64},
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
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
