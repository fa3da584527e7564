"""Pins of the oracle's algebraic recompression (oracle/recompress.py,
restating src/Butterflyalgebra/algrecomp.jl:19-148): its pivoted QR against
LAPACK (scipy's ZGEQP3) and the identities Q^H Q = I and A P = Q R; the
recompressed operator against the dense kernel with the reference's bars
(test/test_algrecomp.jl: assembled at 1e-4, recompressed at 1e-2, error
< 10 x 1e-2, storage not larger); without truncation the operator is
unchanged to rounding."""
import numpy as np
import pytest
import scipy.linalg

from conftest import crandn, two_squares


@pytest.fixture(scope="module")
def orc():
    from oracle import recompress
    return recompress


def test_pqr_matches_lapack_and_is_a_qr(orc):
    rng = np.random.default_rng(5)
    for m, n in ((12, 7), (7, 12), (30, 30)):
        A = crandn(rng, m, n)
        Q, R, p = orc.pqr(A, 0.0)
        _, Rs, ps = scipy.linalg.qr(A, pivoting=True, mode="economic")
        r = min(m, n)
        np.testing.assert_array_equal(p[:r], ps[:r])
        np.testing.assert_allclose(np.abs(np.diag(R)), np.abs(np.diag(Rs)), rtol=1e-10)
        np.testing.assert_allclose(Q.conj().T @ Q, np.eye(r), atol=1e-13)
        assert np.linalg.norm(A[:, p] - Q @ R) <= 1e-13 * np.linalg.norm(A)


def test_pqr_truncation(orc):
    rng = np.random.default_rng(6)
    U, V = crandn(rng, 40, 10), crandn(rng, 10, 25)
    A = (U * np.logspace(0, -9, 10)) @ V
    for rtol in (1e-2, 1e-5):
        Q, R, p = orc.pqr(A, rtol)
        r = Q.shape[1]
        assert 1 <= r < 10 and R.shape == (r, 25)
        # kept pivots above rtol |R11|; the residual is what was dropped
        assert np.all(np.abs(np.diag(R)) > rtol * abs(R[0, 0]))
        assert np.linalg.norm(A[:, p] - Q @ R, 2) <= 30 * rtol * abs(R[0, 0])


def _entries(ex):
    t = ex["table"]
    return int(np.sum(t[t[:, 7] == 0, 5].astype(np.int64) * t[t[:, 7] == 0, 6]))


def test_recompression_error_and_storage(oracle, bfgpu, orc):
    """test_algrecomp.jl:73-166 on the two-squares block."""
    obs, src = two_squares(4096, seed=2)
    K = 2 * np.pi
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    bf = oracle.assemble(To, Ts, 0, 0, K, 1e-4)
    ex = bf.export()
    rng = np.random.default_rng(7)
    x = crandn(rng, 4096)
    rows = np.sort(rng.choice(4096, 256, replace=False))
    A = oracle.helmholtz(K, obs, rows, src, np.arange(4096))
    ye = A @ x
    for tau, bar in ((1e-2, 10 * 1e-2), (1e-5, 10 * 1e-4)):
        rex = orc.recompress(ex, tau)
        assert _entries(rex) <= _entries(ex)
        rb = oracle.import_bf(To, Ts, rex)
        y = rb.matvec(x)
        assert np.linalg.norm(y[rows] - ye) / np.linalg.norm(ye) < bar
    # the coarse recompression shrinks the R levels
    t0, t1 = ex["table"], orc.recompress(ex, 1e-2)["table"]
    R0 = (t0[:, 0] > 0) & (t0[:, 0] < ex["n_stages"] - 1)
    assert _entries({"table": t1[R0]}) < _entries({"table": t0[R0]})
    # no truncation: the same operator to rounding, block keys unchanged
    rex = orc.recompress(ex, 0.0)
    np.testing.assert_array_equal(rex["table"][:, [0, 1, 2, 3, 4, 7, 8, 9]], ex["table"][:, [0, 1, 2, 3, 4, 7, 8, 9]])
    y0 = bf.matvec(x)
    assert np.linalg.norm(oracle.import_bf(To, Ts, rex).matvec(x) - y0) < 1e-12 * np.linalg.norm(y0)
