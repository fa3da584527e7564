"""The learned-rank path against the oracle, record for record.

With LearnedRankEstimator every compression samples (prediction +
oversample) rows; the pivoted QR's adaptive doubling (compressors.jl:134)
and then the residual judge on 8 held-out rows of the same sample prefix
decide whether the sample sufficed, and a rejected compression is redone by
plain adaptive RRQR.  The GPU's predicted guesses are fed to the oracle
(rank_mode 2) with the same judge: compression records (guess, rows, rank),
judge verdicts, skeleton sets and block tables must be identical and FP64
factors agree to 1e-9.  A predictor forced to rank 1 must trigger the
fallback and still give a far field within 10 tau of dense FP64 rows.
"""
import os
import sys

import numpy as np
import pytest

from conftest import crandn, fibonacci_sphere

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


def _far(bfgpu, pts, est, acctype):
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    B = bfgpu.BlockTree(T, T)
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), B, K, tol=TOL,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=acctype,
                                rankestimator=est, farfieldonly=True)
    return op, B


def _compare_sample(op, B, oracle, n_pick, seed):
    """Oracle restatement of factorizations picked at random, fed the GPU's
    guesses; returns (compressions compared, fallbacks among them)."""
    nbf = op.n_factorizations()
    pick = np.sort(np.random.default_rng(seed).choice(nbf, min(nbf, n_pick), replace=False))
    n_comp = n_fb = 0
    for i in pick:
        gex = op.export(int(i))
        no, ns = gex["roots"]
        oex = oracle.assemble(B.testcluster, B.trialcluster, int(no), int(ns), K, TOL, rank_mode=2,
                              fixed=gex["comp"][:, 3].copy(), judge_rows=8, judge_slack=4.0).export()
        np.testing.assert_array_equal(gex["comp"], oex["comp"])
        np.testing.assert_array_equal(gex["judge"], oex["judge"])
        np.testing.assert_array_equal(gex["skel"], oex["skel"])
        np.testing.assert_array_equal(gex["skel_idx"], oex["skel_idx"])
        np.testing.assert_array_equal(gex["table"], oex["table"])
        if len(oex["data"]):
            assert rel(gex["data"], oex["data"]) < 1e-9
        n_comp += len(gex["comp"])
        n_fb += int(np.sum(gex["judge"] == 2))
    return n_comp, n_fb


def _err_vs_dense(op, pts, rows_n, seed):
    from bfgpu import learned
    rng = np.random.default_rng(seed)
    n = pts.shape[0]
    x = crandn(rng, n)
    y = (op @ x.astype(op.eltype)).astype(np.complex128)
    rows = np.sort(rng.choice(n, rows_n, replace=False))
    ye = learned.far_rows(op, pts, K, x, rows)
    return rel(y[rows], ye)


def test_learned_fp64_records_match_oracle(bfgpu, oracle):
    from bfgpu import learned
    net = learned.load(MODEL)
    n = 8192
    pts = fibonacci_sphere(n, 8.0 * np.sqrt(n / 262144))
    say("learned fp64 sphere", n, "with the committed predictor")
    op, B = _far(bfgpu, pts, bfgpu.LearnedRankEstimator(net), np.complex128)
    info = op.info()
    say("compressions", info.n_comp, "judged", info.n_judged, "fallbacks", info.n_fallback,
        "doublings", info.n_retry)
    assert info.n_judged > 0.5 * info.n_comp
    nc, nf = _compare_sample(op, B, oracle, 200, seed=1)
    say("records compared", nc, "fallbacks among them", nf)
    e = _err_vs_dense(op, pts, 48, seed=2)
    say("error vs dense far rows", e)
    assert e < 10 * TOL


def test_learned_host_drawn_samples_match_oracle(bfgpu, oracle, monkeypatch):
    """The judged compressions' samples (sampled rows, then the held-out
    rows of the same prefix) drawn on the host for every prefix longer than
    16 entries: records, verdicts and factors still equal the oracle's."""
    from bfgpu import learned
    net = learned.load(MODEL)
    n = 4096
    pts = fibonacci_sphere(n, 8.0 * np.sqrt(n / 262144))
    monkeypatch.setenv("BFGPU_SAMPLE_DEVICE_MAX", "16")
    say("learned fp64 sphere", n, "host-drawn samples above 16 entries")
    op, B = _far(bfgpu, pts, bfgpu.LearnedRankEstimator(net), np.complex128)
    monkeypatch.delenv("BFGPU_SAMPLE_DEVICE_MAX")
    assert op.info().n_judged > 0
    nc, nf = _compare_sample(op, B, oracle, 120, seed=3)
    say("records compared", nc, "fallbacks among them", nf)


def test_forced_underprediction_falls_back_to_plain_rrqr(bfgpu, oracle):
    from bfgpu import learned
    net = bfgpu.RankPredictor(seed=3)
    net.set_params(np.zeros(net.n_params(), np.float32))  # the network's output layer gives 0 ...
    net.set_norm(np.zeros(6, np.float32), np.ones(6, np.float32), out_shift=1.0, out_scale=1.0)
    assert np.all(net.forward(np.ones((4, 6), np.float32)) == 1.0)  # ... so every predicted rank is 1
    n = 8192
    pts = fibonacci_sphere(n, 8.0 * np.sqrt(n / 262144))
    say("forced rank-1 predictor, no oversampling, fp64 sphere", n)
    op, B = _far(bfgpu, pts, bfgpu.LearnedRankEstimator(net, oversample=0), np.complex128)
    info = op.info()
    say("compressions", info.n_comp, "judged", info.n_judged, "fallbacks", info.n_fallback,
        "doublings", info.n_retry)
    assert info.n_fallback > 0
    nc, nf = _compare_sample(op, B, oracle, 300, seed=4)
    say("records compared", nc, "fallbacks among them", nf)
    assert nf > 0
    e = _err_vs_dense(op, pts, 48, seed=5)
    say("error vs dense far rows with the judge", e)
    assert e < 10 * TOL
    say("same predictor without the judge")
    op0, _ = _far(bfgpu, pts, bfgpu.LearnedRankEstimator(net, oversample=0, judge_rows=0), np.complex128)
    assert op0.info().n_judged == 0 and op0.info().n_fallback == 0
    say("error vs dense far rows without the judge", _err_vs_dense(op0, pts, 48, seed=5))

    say("forced rank-1 predictor, fp32")
    op32, _ = _far(bfgpu, pts, bfgpu.LearnedRankEstimator(net, oversample=0), np.complex64)
    assert op32.info().n_fallback > 0
    e32 = _err_vs_dense(op32, pts, 48, seed=5)
    say("fp32 error vs dense far rows", e32)
    assert e32 < 10 * TOL
