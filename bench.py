"""Benchmark of the MI355X tensor-contraction path (one JSON line on stdout).

Workload: the permuted-layout FP64 contraction
    C[aibj] += 0.5 * A[aebf] * B[fjei],  a=b=e=f=64, i=j=32,
column-major, entries uniform[-1,1], C initialised uniform[-1,1], through
tblis_tensor_mult (FP64 MFMA).  A step is one contraction with inputs
resident in HBM.  Reported:
  value      useful TFLOP/s (2*m*n*k / wall time per step)
  roofline   the GEMM kernel's achieved FLOP/s from HIP events around each
             launch on the library's stream, against the FP64 MFMA peak
  cpu_baseline  the oracle's multithreaded blocked port timed on the host
             cores on a column sample of the same contraction
Extra fields report the other workloads (dense 24^4, DPD, batch) when
--all is given.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "useful TFLOP/s + %MFMA peak; split-BF16 vs FP64-MFMA at stated rel error"
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 matrix (datasheet)


def colmaj(lens):
    s, a = [], 1
    for l in lens:
        s.append(a)
        a *= l
    return s


def permuted_workload(dev, seed=42):
    import torch
    ext = dict(a=64, b=64, e=64, f=64, i=32, j=32)
    lens = lambda s: [ext[c] for c in s]
    lA, lB, lC = lens("aebf"), lens("fjei"), lens("aibj")
    g = torch.Generator(device=dev).manual_seed(seed)
    A = torch.rand(int(np.prod(lA)), dtype=torch.float64, device=dev, generator=g) * 2 - 1
    B = torch.rand(int(np.prod(lB)), dtype=torch.float64, device=dev, generator=g) * 2 - 1
    C = torch.rand(int(np.prod(lC)), dtype=torch.float64, device=dev, generator=g) * 2 - 1
    a = torch.as_strided(A, lA, colmaj(lA))
    b = torch.as_strided(B, lB, colmaj(lB))
    c = torch.as_strided(C, lC, colmaj(lC))
    flops = 2.0 * (64 * 64) * (32 * 32) * (64 * 64)
    nbytes = 8 * (A.numel() + B.numel() + 2 * C.numel())
    return (a, "aebf", b, "fjei", c, "aibj", 0.5, 1.0), flops, nbytes, (lA, lB, lC)


def time_steps(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall / steps * 1e3, float(np.mean(kern)), float(np.min(kern))


def cpu_baseline(lA, lB, lC, sample_cols, threads):
    """Oracle port (multithreaded blocked GEMM over scatter vectors) on the
    first sample_cols columns (n) of the permuted contraction."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, int(np.prod(lA)))
    B = rng.uniform(-1, 1, int(np.prod(lB)))
    C = rng.uniform(-1, 1, int(np.prod(lC)))
    args = (0.5, A, lA, colmaj(lA), "aebf", B, lB, colmaj(lB), "fjei", 1.0, C, lC, colmaj(lC), "aibj")
    po.blis_contract(threads, *args, n_limit=8)  # warm-up
    best, flops = None, 0.0
    for _ in range(2):
        t0 = time.perf_counter()
        flops = po.blis_contract(threads, *args, n_limit=sample_cols)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return flops / best / 1e9, flops, best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-cols", type=int, default=256, help="n columns in the CPU sample")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (reference point for "
                         "FETCH_SIZE/WRITE_SIZE counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see DESIGN.md for how the contraction would shard); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import tblis_amd as T

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    T.set_stream(stream.cuda_stream)
    T.set_async(True)

    wl, flops, nbytes, lens = permuted_workload(dev)
    a, ia, b, ib, c, ic, alpha, beta = wl
    fn = T.prepare_mult(alpha, a, ia, b, ib, beta, c, ic)
    ms_step, ms_kernel, ms_kernel_min = time_steps(fn, args.steps, args.warmup, stream)
    T.set_async(False)
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
    value = flops / (ms_step * 1e-3) / 1e12
    achieved = flops / (ms_kernel * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: uniform[-1,1] from a seeded generator, resident in HBM",
        "config": {"workload": "C[aibj] += 0.5*A[aebf]*B[fjei], a=b=e=f=64, i=j=32, column-major, "
                               "FP64 MFMA (tblis_tensor_mult)",
                   "m": 4096, "n": 1024, "k": 4096, "algorithm": T.last_algorithm()},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64, 4), "traffic": None,
                     "kernel_ms_mean": round(ms_kernel, 4), "kernel_ms_min": round(ms_kernel_min, 4),
                     "algorithmic_bytes": nbytes},
    }
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        gflops, fl, dt = cpu_baseline(*lens, args.cpu_cols, threads)
        out["cpu_baseline"] = {"value": round(gflops / 1e3, 5), "unit": "TFLOP/s", "cores": threads,
                               "kind": "port",
                               "sample": "first %d of 1024 (i,j) columns of the same contraction "
                                         "(%.3g flop, %.2f s), oracle blocked GEMM over scatter "
                                         "vectors, best of 2" % (args.cpu_cols, fl, dt)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
