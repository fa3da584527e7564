"""The benchmark configurations on the GPU.

* The plane-wave ellipsoid (semi-axes 16 x 12 x 8 wavelengths at 1,048,576
  points) reduced to 16,384 points at the same point density (semi-axes
  scaled by sqrt(16384 / 1048576)): learned-rank FP32 factorization with
  the committed predictor, 128 plane-wave right-hand sides on the matrix
  cores in two passes of 64 (bfmatrix.jl:10-50 semantics: column by column
  the single-RHS product), the last level emulating FP32 on the bf16 matrix
  cores (bf16x3).  The FP32 product must match the oracle's FP64 product
  of the same factors to 2e-5 (FP32 matrix cores: 2e-6; the emulated last
  level adds about 5e-6), the plain FP64 GPU operator to 10 tau, and dense
  FP64 far-field rows to 10 tau.
* The sphere far field at its full size (262,144 Fibonacci points, radius
  8 wavelengths): FP32 and FP64 factorizations each within 10 tau of 128
  dense FP64 far-field rows, and within 10 tau of each other.
"""
import os
import sys

import numpy as np
import pytest

from conftest import crandn, ellipsoid_surface, fibonacci_sphere, plane_waves

pytestmark = pytest.mark.gpu

K = 2 * np.pi
TOL = 1e-4
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     "butterflyfactorizations.jl_amd", "bfgpu", "data", "rank_mlp.npz")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_plane_wave_ellipsoid_reduced(bfgpu, oracle):
    from bfgpu import learned
    from bfgpu.planewaves import far_rows_many
    n, nrhs = 16384, 128
    axes = np.array([16.0, 12.0, 8.0]) * np.sqrt(n / 1048576)
    pts = ellipsoid_surface(n, axes, seed=5)
    X = plane_waves(pts, K, nrhs, seed=6)
    net = learned.load(MODEL)
    say("ellipsoid", n, "points, learned fp32 factorization")
    op32, _ = learned.far_operator(pts, acctype=np.complex64, rankestimator=bfgpu.LearnedRankEstimator(net))
    i32 = op32.info()
    say("compressions", i32.n_comp, "judged", i32.n_judged, "fallbacks", i32.n_fallback)
    assert op32.get_option("mfma") >= 1 and 0 < op32.get_option("rhs_lanes") <= nrhs
    say("128 plane waves on the matrix cores")
    Y32 = (op32 @ np.asfortranarray(X.astype(np.complex64))).astype(np.complex128)

    # the oracle's FP64 product of the same (FP32) factors, on columns of
    # both passes
    cols = np.array([0, 1, 31, 63, 64, 65, 100, 127])
    say("oracle product over", op32.n_factorizations(), "factorizations, columns", cols.tolist())
    T = op32.tree.testcluster
    imps = [oracle.import_bf(T, T, op32.export(i)) for i in range(op32.n_factorizations())]
    Yo = oracle.matvec_many(imps, np.asfortranarray(X[:, cols]), nthreads=min(16, os.cpu_count() or 1), rows=n)
    del imps
    e_or = rel(Y32[:, cols], Yo)
    say("fp32 matrix cores (bf16x3 last level) vs oracle of the same factors", e_or)
    assert e_or < 2e-5
    op32.set_option("mfma_emul", 0)
    Y32n = (op32 @ np.asfortranarray(X.astype(np.complex64))).astype(np.complex128)
    op32.set_option("mfma_emul", 1)
    e_or_n = rel(Y32n[:, cols], Yo)
    say("fp32 matrix cores throughout vs oracle", e_or_n)
    assert e_or_n < 2e-6

    say("plain fp64 factorization, 128 right-hand sides")
    op64, _ = learned.far_operator(pts, acctype=np.complex128)
    Y64 = op64 @ X
    e64 = rel(Y32, Y64)
    say("learned fp32 vs plain fp64 (all 128)", e64)
    assert e64 < 10 * TOL
    rows = np.sort(np.random.default_rng(7).choice(n, 24, replace=False))
    Ye = far_rows_many(op64, pts, K, X, rows)
    e_d32, e_d64 = rel(Y32[rows], Ye), rel(Y64[rows], Ye)
    say("vs dense far rows: fp32", e_d32, "fp64", e_d64)
    assert e_d32 < 10 * TOL and e_d64 < 10 * TOL


def test_sphere_far_field_full_size(bfgpu):
    from bfgpu import learned
    n = 262144
    pts = fibonacci_sphere(n, 8.0)
    x = crandn(np.random.default_rng(11), n)
    rows = np.sort(np.random.default_rng(12).choice(n, 128, replace=False))
    say("sphere", n, "fp32 factorization")
    op32, t32 = learned.far_operator(pts, acctype=np.complex64)
    say("factorized in", round(t32, 2), "s;", op32.n_factorizations(), "factorizations")
    y32 = (op32 @ x.astype(np.complex64)).astype(np.complex128)
    ye = learned.far_rows(op32, pts, K, x, rows)
    del op32
    say("sphere", n, "fp64 factorization")
    op64, t64 = learned.far_operator(pts, acctype=np.complex128)
    y64 = op64 @ x
    del op64
    e32, e64, e3264 = rel(y32[rows], ye), rel(y64[rows], ye), rel(y32, y64)
    say("vs 128 dense far rows: fp32", e32, "fp64", e64, "; fp32 vs fp64", e3264)
    assert e32 < 10 * TOL and e64 < 10 * TOL and e3264 < 10 * TOL
