"""Far field of an ellipsoid under plane-wave excitation: learned-rank FP32
factorization, a multi-RHS product on the matrix cores, the HBM footprint,
and the error against FP64.

Points are uniform in area on the ellipsoid (octree leaf 64, admissibility
2.0, k = 2 pi, tolerance 1e-4); the right-hand sides are plane waves
exp(-i k d . x) with directions d uniform on the sphere.  Errors are taken
against an FP64 product computed two ways: dense FP64 far-field rows on the
host (a sample of rows, every right-hand side), and, when the device has
room for it after the FP32 operator is freed, the plain FP64 GPU
factorization applied to the first few right-hand sides.

    python -m bfgpu.planewaves --points 1048576 --axes 16 12 8 --rhs 128 \\
        --model bfgpu/data/rank_mlp.npz [--near]

With --near the operator's near field is assembled too (FP32) and applied to
the same plane waves.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import BfgpuError, LearnedRankEstimator, default_context
from .geometry import ellipsoid_surface, plane_waves
from .learned import K, far_operator, load


def _host_gb():
    try:
        import psutil
        return psutil.Process().memory_info().rss / 1e9
    except ImportError:
        return float("nan")


def far_rows_many(op, points, k, X, rows):
    """Dense FP64 far-field rows of the operator applied to every column of
    X (farmatrix, matrixstructs.jl:50-61): for each row, the sum over the
    sources of the admissible pairs whose observer node holds it."""
    T = op.tree.testcluster
    S = op.tree.trialcluster
    by_obs = {}
    for o, s in op.farinteractions:
        by_obs.setdefault(int(o), []).append(int(s))
    leaf_of = np.empty(points.shape[0], np.int64)
    for v in T.leaves():
        leaf_of[T.cluster_values(v)] = v
    out = np.zeros((len(rows), X.shape[1]), complex)
    for j, r in enumerate(rows):
        mask = np.zeros(points.shape[0], bool)
        node = int(leaf_of[r])
        while node >= 0:
            for s in by_obs.get(node, ()):
                mask[S.cluster_values(s)] = True
            node = T.cluster_parent(node)
        d = np.linalg.norm(points[mask] - points[r], axis=1)
        out[j] = (np.exp(-1j * k * d) / (4 * np.pi * d)) @ X[mask]
    return out


def near_field(op, n_rhs, Xd, reps, ctx, log):
    """The operator's near field (BlockSparseNear, FP32) assembled beside
    the far field and applied to the same right-hand sides."""
    import torch
    from . import BlockSparseNear, HelmholtzKernelMatrix
    t0 = time.perf_counter()
    N = BlockSparseNear(HelmholtzKernelMatrix(K), op.tree, op.nearpairs, acctype=np.complex64, ctx=ctx)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    inf = N.info()
    Yd = torch.zeros(n_rhs, Xd.shape[0], dtype=torch.complex64, device="cuda").t()
    N.mul(Yd, Xd)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        N.mul(Yd, Xd)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = 1e3 * min(times)
    log(f"near field: {inf['pairs']} pairs, {inf['entries'] * 8 / 1e9:.1f} GB, assembled in {t_asm:.2f} s, "
        f"{n_rhs}-RHS product {ms:.1f} ms")
    out = {"pairs": inf["pairs"], "entries": inf["entries"], "device_bytes": inf["device_bytes"],
           "assemble_s": round(t_asm, 2), "matvec_ms": round(ms, 2)}
    del N, Yd
    torch.cuda.empty_cache()
    return out


def run(points=1048576, axes=(16.0, 12.0, 8.0), n_rhs=128, model=None, err_rows=16, fp64_rhs=8, reps=3,
        seed=0, ctx=None, log=print, near=False):
    import torch
    ctx = ctx or default_context()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    pts = ellipsoid_surface(points, axes, seed=seed)
    X = plane_waves(pts, K, n_rhs, seed=seed + 1)
    free0, total = torch.cuda.mem_get_info()
    net = load(model, ctx=ctx) if model else None
    log(f"{points} points on the {axes} ellipsoid, {n_rhs} plane waves; factorizing "
        f"({'learned' if net else 'plain'} ranks, FP32)")
    op, t_fact = far_operator(pts, acctype=np.complex64, ctx=ctx,
                              rankestimator=LearnedRankEstimator(net, 10) if net else None)
    inf = op.info()
    free1, _ = torch.cuda.mem_get_info()
    rss = _host_gb()
    log(f"factorized in {t_fact:.1f} s: {len(op.farinteractions)} admissible pairs, "
        f"{inf.factor_bytes / 1e9:.1f} GB of factors, host memory {rss:.1f} GB")
    Xd = torch.from_numpy(X.astype(np.complex64).T.copy()).cuda().t()  # column-major on the device
    Yd = torch.zeros(n_rhs, points, dtype=torch.complex64, device="cuda").t()
    op.mul(Yd, Xd)
    torch.cuda.synchronize()
    free2, _ = torch.cuda.mem_get_info()
    held = op.info().device_bytes
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        op.mul(Yd, Xd)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    Y = Yd.cpu().numpy()
    ms = 1e3 * min(times)
    log(f"{n_rhs}-RHS matvec {ms:.1f} ms (best of {reps}); dense FP64 rows next")
    rng = np.random.default_rng(seed + 2)
    rows = np.sort(rng.choice(points, err_rows, replace=False))
    ye = far_rows_many(op, pts, K, X, rows)
    rec, _ = op.compressions()
    out = {"points": points, "axes": list(axes), "rhs": n_rhs, "tol": 1e-4,
           "ranks": "learned" if net else "plain",
           "factorize_s": round(t_fact, 2), "matvec_ms": round(ms, 2), "matvec_ms_per_rhs": round(ms / n_rhs, 3),
           "admissible_pairs": int(len(op.farinteractions)), "compressions": int(inf.n_comp),
           "doublings": int(inf.n_retry), "judged": int(inf.n_judged), "fallbacks": int(inf.n_fallback), "mean_rows": float(rec[:, 5].mean()),
           "mean_rank": float(rec[:, 6].mean()), "factor_bytes": int(inf.factor_bytes),
           "operator_device_bytes_after_factorization": int(inf.device_bytes),
           "operator_device_bytes_after_matvec": int(held),
           "device_in_use_after_factorization_gb": round((free0 - free1) / 1e9, 2),
           "device_in_use_after_matvec_gb": round((free0 - free2) / 1e9, 2), "hbm_total_gb": round(total / 1e9, 1),
           "slot_bytes_per_rhs": int(inf.slot_len) * 8, "host_rss_gb_after_factorization": round(rss, 1),
           "err_vs_dense_rows": rel(Y[rows].astype(complex), ye), "dense_rows": int(err_rows)}
    if near:
        out["near_field"] = near_field(op, n_rhs, Xd, reps, ctx, log)
        out["whole_operator_matvec_ms"] = round(ms + out["near_field"]["matvec_ms"], 2)
    del op, Xd, Yd
    torch.cuda.empty_cache()
    free3, _ = torch.cuda.mem_get_info()
    log(f"error against {err_rows} dense FP64 rows: {out['err_vs_dense_rows']:.2e}; "
        f"{(free0 - free3) / 1e9:.1f} GB still in use after freeing the operator")
    if fp64_rhs > 0:
        log(f"plain FP64 GPU factorization for the first {fp64_rhs} right-hand sides")
        try:
            op64, t64 = far_operator(pts, acctype=np.complex128, ctx=ctx)
        except BfgpuError as e:  # device memory: reported, not fatal
            out["fp64_gpu"] = f"not run: {e}"
        else:
            X64 = torch.from_numpy(np.ascontiguousarray(X[:, :fp64_rhs].T)).cuda().t()
            Y64 = torch.zeros(fp64_rhs, points, dtype=torch.complex128, device="cuda").t()
            op64.mul(Y64, X64)
            y64 = Y64.cpu().numpy()
            out["fp64_gpu"] = {"factorize_s": round(t64, 2), "factor_bytes": int(op64.info().factor_bytes),
                               "rhs": fp64_rhs, "err_fp32_learned_vs_fp64_plain":
                                   rel(Y[:, :fp64_rhs].astype(complex), y64),
                               "err_fp64_vs_dense_rows": rel(y64[rows], ye[:, :fp64_rhs])}
            del op64, X64, Y64
            torch.cuda.empty_cache()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m bfgpu.planewaves")
    ap.add_argument("--points", type=int, default=1048576)
    ap.add_argument("--axes", type=float, nargs=3, default=(16.0, 12.0, 8.0))
    ap.add_argument("--rhs", type=int, default=128)
    ap.add_argument("--model", default=None, help="rank predictor (.npz); without it the ranks are plain")
    ap.add_argument("--err-rows", type=int, default=16)
    ap.add_argument("--fp64-rhs", type=int, default=8, help="right-hand sides of the FP64 GPU check (0: none)")
    ap.add_argument("--near", action="store_true", help="also assemble and apply the near field (FP32)")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    print(json.dumps(run(a.points, tuple(a.axes), a.rhs, a.model, a.err_rows, a.fp64_rhs, log=log, near=a.near)),
          flush=True)


if __name__ == "__main__":
    main()
