"""Benchmark: butterfly matvec time and HBM roofline fraction on one MI355X.

Workload (one step = one matvec over one right-hand side): the single
admissible block between two parallel 16 x 16 wavelength squares --
65,536 uniform-random observers on [0,16]^2 x {0} and 65,536 sources on
[0,16]^2 x {32}, Helmholtz kernel exp(-ik|x-y|)/(4 pi |x-y|), k = 2 pi,
balanced quadtree with 64-point leaves (6 levels), tolerance 1e-4.  The
factorization is computed on the GPU in complex FP32 (the measured path) and
complex FP64 (the plain path); the timed region applies the FP32
factorization to a complex-Gaussian vector already resident in HBM.

Prints one JSON line.  --gpus N > 1 is rejected: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "butterflyfactorizations.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(n, seed=0):
    side = 16.0
    rng = np.random.default_rng(seed)
    obs = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.zeros(n)]
    src = np.c_[rng.uniform(0, side, n), rng.uniform(0, side, n), np.full(n, 32.0)]
    return obs, src


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    import torch

    import bfgpu

    k, n, tol = 2 * np.pi, args.n, args.tol
    obs, src = workload(n)
    bt = bfgpu.BlockTree(bfgpu.ClusterTree(obs, dims=2, leaf_size=64),
                         bfgpu.ClusterTree(src, dims=2, leaf_size=64))
    ctx = bfgpu.Context(0)
    torch.cuda.init()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    km = bfgpu.HelmholtzKernelMatrix(k)

    log("factorize fp32")
    t0 = time.perf_counter()
    g32 = bfgpu.assemble_BF(km, bt, 0, 0, k, tol, acctype=np.complex64, ctx=ctx)
    t_f32 = time.perf_counter() - t0
    log(f"factorize fp32 {t_f32:.3f}s; fp64")
    t0 = time.perf_counter()
    g64 = bfgpu.assemble_BF(km, bt, 0, 0, k, tol, acctype=np.complex128, ctx=ctx)
    t_f64 = time.perf_counter() - t0
    log(f"factorize fp64 {t_f64:.3f}s")
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)

    gen = torch.Generator(device="cpu").manual_seed(0)
    xr = torch.randn(n, generator=gen, dtype=torch.float64) / np.sqrt(2)
    xi = torch.randn(n, generator=gen, dtype=torch.float64) / np.sqrt(2)
    x64 = torch.complex(xr, xi).cuda()
    x32 = x64.to(torch.complex64)
    y32 = torch.zeros(n, dtype=torch.complex64, device="cuda")
    y64 = torch.zeros(n, dtype=torch.complex128, device="cuda")

    def timed(g, x, y, steps, warm):
        for _ in range(warm):
            g.mul(y, x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.mul(y, x)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e3

    ms32 = timed(g32, x32, y32, args.steps, args.warmup)
    ms64 = timed(g64, x64, y64, args.steps, args.warmup)
    log(f"matvec fp32 {ms32:.4f} ms, fp64 {ms64:.4f} ms")

    # multi-RHS (64 right-hand sides) in FP32
    X = torch.complex(torch.randn(64, n, generator=gen), torch.randn(64, n, generator=gen)).cuda()
    X = (X / np.sqrt(2)).to(torch.complex64).t()
    Y = torch.zeros(64, n, dtype=torch.complex64, device="cuda").t()
    ms32_64 = timed(g32, X, Y, max(args.steps // 5, 3), 2)

    # per-stage HIP-event timings (context stream = torch's current stream)
    st_ms, st_bytes, tot_ms = g32.timed(x32, y32, iters=max(args.steps, 10))
    dom = int(np.argmax(st_ms))
    achieved = st_bytes[dom] / (st_ms[dom] * 1e-3) / 1e9
    whole = sum(st_bytes) / (ms32 * 1e-3) / 1e9

    # errors against 512 sampled dense rows (FP64 oracle kernel on the host)
    from oracle import oracle
    rng = np.random.default_rng(1)
    rows = np.sort(rng.choice(n, 512, replace=False))
    A = oracle.helmholtz(k, obs, rows, src, np.arange(n))
    xh = x64.cpu().numpy()
    ye = A @ xh
    e32 = float(np.linalg.norm(y32.cpu().numpy()[rows] - ye) / np.linalg.norm(ye))
    e64 = float(np.linalg.norm(y64.cpu().numpy()[rows] - ye) / np.linalg.norm(ye))
    i32 = g32.info()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle FP64 flat matvec, 1 thread)")
        ob = oracle.import_bf(bt.testcluster, bt.trialcluster, g64.export())
        ts = []
        t_end = time.perf_counter() + 20.0
        while len(ts) < 20 or (time.perf_counter() < t_end and len(ts) < 200):
            t0 = time.perf_counter()
            ob.matvec(xh)
            ts.append((time.perf_counter() - t0) * 1e3)
            if len(ts) >= 20 and time.perf_counter() > t_end:
                break
        cpu = {"value": float(np.min(ts)), "unit": "ms", "cores": 1, "kind": "port",
               "median_ms": float(np.median(ts)),
               "sample": f"{len(ts)} full FP64 matvecs of the same factorization (flat-workspace "
                         "mul!, bfvector.jl:200-296, restated in C), min reported"}

    out = {
        "metric": "butterfly matvec ms & %HBM roofline, FP32-batched vs FP64 plain at 1e-4 err",
        "value": round(ms32, 5), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms32, 5), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"two parallel squares, {n} points each, quadtree leaf 64, "
                               f"tol {tol:g}, k=2pi, 1 RHS, FP32 factorization", "n_points": n,
                   "levels": int(g32.info().max_stages) - 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"stage {dom}", "stage_ms": [round(v, 5) for v in st_ms],
                     "stage_bytes": st_bytes, "whole_matvec_gbs": round(whole, 1)},
        "cpu_baseline": cpu,
        "fp64_ms": round(ms64, 5), "fp32_64rhs_ms": round(ms32_64, 5),
        "factorize_s": {"fp32": round(t_f32, 3), "fp64": round(t_f64, 3)},
        "err_vs_dense_512rows": {"fp32": e32, "fp64": e64},
        "factor_bytes": {"fp32": int(i32.factor_bytes), "fp64": int(g64.info().factor_bytes)},
        "compressions": int(i32.n_comp), "adaptive_retries": int(i32.n_retry),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
