"""Learned rank predictor end to end on the GPU: rank logs of plain
adaptive RRQR on seeded ellipsoids -> MLP trained by the library's kernels
-> far-field factorization with predicted ranks.  Bars (written below): the
rates of adaptive doubling and of the residual judge's fallback to plain
RRQR each stay under 5 %, the row samples shrink by
at least 10 % against the butterfly estimator, and the learned far field
stays within 10 tol of dense FP64 rows and of the plain factorization."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_learned_ranks_shrink_samples_without_losing_accuracy(bfgpu):
    from bfgpu import learned
    tol = 1e-4
    specs = learned.ellipsoid_specs(6, seed=3, axes=(1.0, 3.0), points=(4096, 8192))
    say("collect rank logs on", len(specs), "ellipsoids")
    f, r = learned.collect(specs, per_geometry=20000, log=say)
    assert f.shape == (len(r), 6) and np.all(r >= 1)
    net, loss = learned.train(f, r, epochs=30, seed=1)
    pred = net.forward(f)
    say("train loss", loss, "mean |rank error|", float(np.mean(np.abs(pred - r))))
    assert np.mean(np.abs(pred - r)) < 1.5
    res = learned.evaluate(net, points=16384, tol=tol, err_rows=32)
    P, Lr = res["plain"], res["learned"]
    say("plain", P)
    say("learned", Lr)
    say("learned vs plain", res["learned_vs_plain"])
    assert Lr["fallback_rate"] <= 0.05 and Lr["doubling_rate"] <= 0.05
    assert Lr["judged"] > 0
    assert Lr["mean_rows"] < 0.9 * P["mean_rows"]
    assert P["err_vs_dense"] < 10 * tol and Lr["err_vs_dense"] < 10 * tol
    assert res["learned_vs_plain"] < 10 * tol


def test_predictor_save_load_roundtrip(bfgpu, tmp_path):
    from bfgpu import learned
    rng = np.random.default_rng(0)
    f = rng.standard_normal((512, 6)).astype(np.float32)
    r = (3 + f[:, 0] ** 2).astype(np.float32)
    net, _ = learned.train(f, r, epochs=2, batch=128)
    learned.save(net, tmp_path / "m.npz")
    back = learned.load(tmp_path / "m.npz")
    assert np.array_equal(back.get_params(), net.get_params())
    assert np.allclose(back.forward(f), net.forward(f), rtol=0, atol=1e-5)
