"""Learned rank predictor: training data from the method's own rank logs,
training and inference on the GPU, and its evaluation against plain
adaptive RRQR.

The reference calibrates its rank estimators from logged compressions
(evaluation/ranklogging.jl:16-61: `RankLogger` records per compression,
compressors.jl:52-69, then a least-squares fit).  Here every compression of
a far-field factorization exports six geometric features and its rank
(`ButterflyFactorization.compressions`); an MLP 6 -> 128 -> 128 -> 128 -> 1
(ReLU, FP32) is trained on them by the library's own forward, backward and
Adam kernels.  With `LearnedRankEstimator` the prediction plus 10 rows sizes
the row sample of PartialQR (compressors.jl:99-101); the adaptive doubling
of plain RRQR (compressors.jl:134-151) still applies when the sampled block
turns out nearly full rank, and a residual judge on 8 held-out rows of the
same sample sends every compression whose interpolative decomposition
misses the tolerance there back to plain adaptive RRQR.

    python -m bfgpu.learned train --geometries 200 --out rank_mlp.npz
    python -m bfgpu.learned evaluate --model rank_mlp.npz --points 65536
"""
import argparse
import json
import sys
import time

import numpy as np

from . import (BlockTree, CenterDistanceAdmissibility, ClusterTree, HelmholtzKernelMatrix,
               LearnedRankEstimator, PetrovGalerkinBF, RankPredictor, default_context)
from .geometry import ellipsoid_surface, fibonacci_sphere

K = 2 * np.pi


def far_operator(points, k=K, tol=1e-4, acctype=np.complex64, rankestimator=None, ctx=None, leaf=64,
                 separation=2.0):
    """The far field of the Helmholtz operator on a point cloud (octree with
    at most `leaf` points per leaf, centre-distance admissibility)."""
    T = ClusterTree(points, dims=3, leaf_size=leaf, mode="maxpoints")
    bt = BlockTree(T, T)
    t0 = time.perf_counter()
    op = PetrovGalerkinBF(HelmholtzKernelMatrix(k), bt, k, tol=tol,
                          admissibility=CenterDistanceAdmissibility(separation), acctype=acctype,
                          rankestimator=rankestimator, farfieldonly=True, ctx=ctx)
    return op, time.perf_counter() - t0


def ellipsoid_specs(count, seed=0, axes=(2.0, 10.0), points=(16384, 65536)):
    """`count` seeded random ellipsoids: (semi-axes in wavelengths, points,
    point seed)."""
    rng = np.random.default_rng(seed)
    return [(tuple(float(a) for a in rng.uniform(axes[0], axes[1], 3)),
             int(rng.integers(points[0], points[1] + 1)), int(seed) * 100003 + i) for i in range(count)]


def collect(specs, per_geometry=20000, seed=0, log=print, **kw):
    """Rank logs of plain adaptive RRQR over the geometries: features
    (n, 6) float32 and ranks (n,) float32, at most `per_geometry` records
    drawn at random from each geometry's compressions."""
    rng = np.random.default_rng(seed)
    feats, ranks = [], []
    for j, (axes, n, gseed) in enumerate(specs):
        pts = ellipsoid_surface(n, axes, seed=gseed)
        op, t = far_operator(pts, **kw)
        rec, feat = op.compressions()
        keep = np.arange(len(rec))
        if per_geometry and len(rec) > per_geometry:
            keep = np.sort(rng.choice(len(rec), per_geometry, replace=False))
        feats.append(feat[keep])
        ranks.append(rec[keep, 6].astype(np.float32))
        log(f"geometry {j}: semi-axes {tuple(round(a, 2) for a in axes)}, {n} points, "
            f"{len(rec)} compressions, mean rank {rec[:, 6].mean():.2f}, {t:.2f} s")
        del op
    return np.concatenate(feats), np.concatenate(ranks)


def train(features, ranks, hidden=128, layers=4, epochs=40, batch=1024, lr=1e-3, seed=0, ctx=None):
    """Fit a RankPredictor (inputs standardised, output mapped back to a
    rank by the training mean and spread).  Returns (predictor, loss)."""
    f = np.ascontiguousarray(features, np.float32)
    r = np.ascontiguousarray(ranks, np.float32)
    mu = f.mean(axis=0)
    sd = f.std(axis=0)
    sd = np.where(sd > 1e-6, sd, 1.0).astype(np.float32)
    net = RankPredictor(f.shape[1], hidden, layers, seed=seed, ctx=ctx)
    net.set_norm(mu, sd, float(r.mean()), float(max(r.std(), 1e-3)))
    loss = net.train(f, r, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, loss


def save(net, path):
    mu, sd, shift, scale = net.norm
    np.savez(path, params=net.get_params(), mu=mu, sigma=sd, shift=np.float32(shift), scale=np.float32(scale),
             shape=np.array([net.n_in, net.hidden, net.n_layers], np.int32))


def load(path, ctx=None):
    d = np.load(path, allow_pickle=False)
    n_in, hidden, layers = (int(v) for v in d["shape"])
    net = RankPredictor(n_in, hidden, layers, ctx=ctx)
    net.set_params(d["params"])
    net.set_norm(d["mu"], d["sigma"], float(d["shift"]), float(d["scale"]))
    return net


def far_rows(op, points, k, x, rows):
    """Dense FP64 far-field rows (the sum over the admissible pairs whose
    observer node holds the row, farmatrix, matrixstructs.jl:50-61),
    computed on the host for error reporting."""
    T = op.tree.testcluster
    S = op.tree.trialcluster
    by_obs = {}
    for o, s in op.farinteractions:
        by_obs.setdefault(int(o), []).append(int(s))
    leaf_of = np.empty(points.shape[0], np.int64)
    for v in T.leaves():
        leaf_of[T.cluster_values(v)] = v
    out = np.zeros(len(rows), complex)
    for j, r in enumerate(rows):
        mask = np.zeros(points.shape[0], bool)
        node = int(leaf_of[r])
        while node >= 0:
            for s in by_obs.get(node, ()):
                mask[S.cluster_values(s)] = True
            node = T.cluster_parent(node)
        d = np.linalg.norm(points[mask] - points[r], axis=1)
        out[j] = np.sum(np.exp(-1j * k * d) / (4 * np.pi * d) * x[mask])
    return out


def summary(op, t_wall):
    i = op.info()
    rec, _ = op.compressions()
    return {"factorize_s": round(t_wall, 3), "gpu_fill_qr_s": round(i.t_qr_ms * 1e-3, 3),
            "gpu_interp_s": round(i.t_interp_ms * 1e-3, 3), "sample_prep_s": round(i.t_sample_ms * 1e-3, 3),
            "compressions": int(i.n_comp), "doublings": int(i.n_retry),
            "doubling_rate": float(i.n_retry) / max(int(i.n_comp), 1),
            "judged": int(i.n_judged), "fallbacks": int(i.n_fallback),
            "fallback_rate": float(i.n_fallback) / max(int(i.n_comp), 1),
            "mean_rows": float(rec[:, 5].mean()), "mean_rank": float(rec[:, 6].mean()),
            "factor_bytes": int(i.factor_bytes)}


def evaluate(net, points=65536, tol=1e-4, err_rows=64, seed=0, ctx=None, acctype=np.complex64):
    """Plain adaptive RRQR against learned ranks on the sphere (radius
    8 sqrt(points / 262144) wavelengths): factorization time and its GPU
    part, sampled rows, fallback rate, and the error of each against dense
    FP64 far-field rows and against each other."""
    pts = fibonacci_sphere(points, 8.0 * np.sqrt(points / 262144))
    plain, t_plain = far_operator(pts, tol=tol, acctype=acctype, ctx=ctx)
    learned, t_learned = far_operator(pts, tol=tol, acctype=acctype, ctx=ctx,
                                      rankestimator=LearnedRankEstimator(net, 10))
    rng = np.random.default_rng(seed)
    x = ((rng.standard_normal(points) + 1j * rng.standard_normal(points)) / np.sqrt(2)).astype(acctype)
    yp = (plain @ x).astype(np.complex128)
    yl = (learned @ x).astype(np.complex128)
    rows = np.sort(rng.choice(points, err_rows, replace=False))
    ye = far_rows(plain, pts, K, x.astype(np.complex128), rows)
    err = lambda y: float(np.linalg.norm(y[rows] - ye) / np.linalg.norm(ye))
    return {"points": points, "tol": tol,
            "plain": dict(summary(plain, t_plain), err_vs_dense=err(yp)),
            "learned": dict(summary(learned, t_learned), err_vs_dense=err(yl)),
            "learned_vs_plain": float(np.linalg.norm(yl - yp) / np.linalg.norm(yp))}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m bfgpu.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--geometries", type=int, default=200)
    t.add_argument("--min-points", type=int, default=16384)
    t.add_argument("--max-points", type=int, default=65536)
    t.add_argument("--per-geometry", type=int, default=20000)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", required=True)
    e.add_argument("--points", type=int, default=65536)
    e.add_argument("--err-rows", type=int, default=64)
    a = ap.parse_args(argv)
    ctx = default_context()
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    if a.cmd == "train":
        specs = ellipsoid_specs(a.geometries, a.seed, points=(a.min_points, a.max_points))
        t0 = time.perf_counter()
        f, r = collect(specs, a.per_geometry, a.seed, log=log, ctx=ctx)
        t1 = time.perf_counter()
        net, loss = train(f, r, epochs=a.epochs, seed=a.seed, ctx=ctx)
        t2 = time.perf_counter()
        save(net, a.out)
        pred = net.forward(f[:200000])
        print(json.dumps({"geometries": a.geometries, "samples": int(len(r)), "collect_s": round(t1 - t0, 1),
                          "train_s": round(t2 - t1, 2), "final_loss": loss,
                          "mean_abs_rank_error": float(np.mean(np.abs(pred - r[:200000]))),
                          "under_predicted_by_more_than_2": float(np.mean(pred < r[:200000] - 2))}))
    else:
        net = load(a.model, ctx=ctx)
        print(json.dumps(evaluate(net, a.points, err_rows=a.err_rows, ctx=ctx)))


if __name__ == "__main__":
    main()
