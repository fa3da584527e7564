"""FP64 plain adaptive RRQR against FP32 learned ranks on the far field of
a sphere: factorization time, matvec time, factor bytes and the error of
each against dense FP64 far-field rows.

Both forms run on the GPU.  "Plain" is the reference's default path
(butterfly rank estimator, adaptive PartialQR, complex FP64 factors and
products; compressors.jl:81-165, rankestimation.jl:32-40); "learned" sizes
the row sample from the rank predictor and stores and applies the factors
in complex FP32 (FP64 only for the QR column norms and the triangular
solves).  Each form's matvec is autotuned before it is timed; each form is
factorized several times, interleaved with the other, and its fastest run
counts.

    python -m bfgpu.compare --points 262144 --model bfgpu/data/rank_mlp.npz
"""
import argparse
import json
import sys
import time

import numpy as np

from . import LearnedRankEstimator, default_context
from .geometry import fibonacci_sphere
from .learned import K, far_operator, far_rows, load


def _factorize_time(pts, acctype, rankestimator, ctx):
    op, t = far_operator(pts, acctype=acctype, rankestimator=rankestimator, ctx=ctx)
    del op
    return t


def _form(pts, acctype, rankestimator, x, rows, ref, iters, ctx, log, t_prev=()):
    import torch
    op, t_fact = far_operator(pts, acctype=acctype, rankestimator=rankestimator, ctx=ctx)
    t_fact = min((t_fact,) + tuple(t_prev))
    inf = op.info()
    if "rows" not in ref:  # dense FP64 far-field rows, from the first form's far list
        ref["rows"] = far_rows(op, pts, K, x, rows)
    ye = ref["rows"]
    tw = torch.complex128 if acctype == np.complex128 else torch.complex64
    xd = torch.from_numpy(x.astype(acctype)).cuda()
    yd = torch.zeros(len(x), dtype=tw, device="cuda")
    t0 = time.perf_counter()
    op.autotune(nrhs=1, iters=3)
    t_tune = time.perf_counter() - t0
    op.mul(yd, xd)
    torch.cuda.synchronize()
    _, _, ms = op.timed(xd, yd, iters=iters)
    op.mul(yd, xd)
    y = yd.cpu().numpy().astype(np.complex128)
    err = float(np.linalg.norm(y[rows] - ye) / np.linalg.norm(ye))
    rec, _ = op.compressions()
    out = {"factorize_s": round(t_fact, 3), "autotune_s": round(t_tune, 2), "matvec_ms": round(ms, 4),
           "factorize_plus_matvec_s": round(t_fact + ms * 1e-3, 3),
           "gpu_fill_qr_s": round(inf.t_qr_ms * 1e-3, 3), "gpu_interp_s": round(inf.t_interp_ms * 1e-3, 3),
           "mean_rows": float(rec[:, 5].mean()), "mean_rank": float(rec[:, 6].mean()),
           "doublings": int(inf.n_retry), "judged": int(inf.n_judged), "fallbacks": int(inf.n_fallback), "factor_bytes": int(inf.factor_bytes), "err_vs_dense": err}
    log(f"{np.dtype(acctype).name} {'learned' if rankestimator else 'plain'}: {out}")
    del op, xd, yd
    torch.cuda.empty_cache()
    return out, y


def compare(points=262144, model=None, err_rows=32, iters=20, seed=0, repeats=2, ctx=None, log=print):
    """Return the comparison record (one dict per form plus ratios)."""
    ctx = ctx or default_context()
    pts = fibonacci_sphere(points, 8.0 * np.sqrt(points / 262144))
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(points) + 1j * rng.standard_normal(points)) / np.sqrt(2)
    rows = np.sort(rng.choice(points, err_rows, replace=False))
    # a small operator first, so that library loading and device start-up
    # are not charged to the first form's factorization; then each form is
    # factorized `repeats` times, interleaved, and its fastest run counts
    # (the first large factorization of a process also pays for first-touch
    # host memory)
    far_operator(fibonacci_sphere(4096, 1.0), ctx=ctx)
    net = load(model, ctx=ctx) if model else None
    est = LearnedRankEstimator(net, 10) if net else None
    t64, t32 = [], []
    for _ in range(repeats - 1):
        t64.append(_factorize_time(pts, np.complex128, None, ctx))
        t32.append(_factorize_time(pts, np.complex64, est, ctx))
        log(f"factorization times so far: FP64 plain {t64}, FP32 {t32}")
    ref = {}
    plain, yp = _form(pts, np.complex128, None, x, rows, ref, iters, ctx, log, t64)
    learned, yl = _form(pts, np.complex64, est, x, rows, ref, iters, ctx, log, t32)
    return {"points": points, "tol": 1e-4, "err_rows": err_rows,
            "fp64_plain": plain, ("fp32_learned" if net else "fp32_plain"): learned,
            "factorize_speedup": round(plain["factorize_s"] / learned["factorize_s"], 3),
            "matvec_speedup": round(plain["matvec_ms"] / learned["matvec_ms"], 3),
            "factorize_plus_matvec_speedup": round(plain["factorize_plus_matvec_s"]
                                                   / learned["factorize_plus_matvec_s"], 3),
            "fp32_vs_fp64_matvec": float(np.linalg.norm(yl - yp) / np.linalg.norm(yp))}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m bfgpu.compare")
    ap.add_argument("--points", type=int, default=262144)
    ap.add_argument("--model", default=None, help="rank predictor (.npz); without it the FP32 form is plain")
    ap.add_argument("--err-rows", type=int, default=32)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--repeats", type=int, default=2, help="factorizations per form (fastest counts)")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    print(json.dumps(compare(a.points, a.model, a.err_rows, a.iters, repeats=a.repeats, log=log)), flush=True)


if __name__ == "__main__":
    main()
