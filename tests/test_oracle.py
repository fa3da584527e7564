"""Pins of the CPU FP64 oracle (oracle/bfo.c) against LAPACK fixtures, closed
forms, invariants and dense products, mirroring the reference's tolerance
tests (test/test_subroutines.jl:174-182, test/test_bfadjtr.jl:133-185)."""
import os

import numpy as np
import pytest

from conftest import crandn, two_squares

GOLD = os.path.join(os.path.dirname(__file__), "golden", "cpqr_zgeqp3.npz")


# ---------------------------------------------------------------- pivoted QR
def test_cpqr_matches_lapack_zgeqp3(oracle):
    g = np.load(GOLD)
    for i in range(int(g["n"])):
        A = g[f"A{i}"]
        r, piv, R, tau = oracle.cpqr(A, 0.0)
        mn = min(A.shape)
        assert r == mn
        np.testing.assert_array_equal(piv[:mn], g[f"jpvt{i}"][:mn])
        rd = np.abs(np.diag(R[:mn, :mn]))
        np.testing.assert_allclose(rd, g[f"rdiag{i}"][:mn], rtol=1e-10, atol=1e-13 * rd[0])
        # |R| equals LAPACK's up to the column phases
        np.testing.assert_allclose(np.abs(np.triu(R[:mn, :])), np.abs(g[f"R{i}"]),
                                   rtol=1e-8, atol=1e-12 * rd[0])


def test_cpqr_reconstructs_and_truncates(oracle):
    rng = np.random.default_rng(3)
    m, n, r0 = 40, 60, 12
    U = crandn(rng, m, r0)
    V = crandn(rng, r0, n)
    s = np.logspace(0, -10, r0)
    A = (U * s) @ V
    for rtol in (1e-3, 1e-6, 1e-9):
        r, piv, R, tau = oracle.cpqr(A, rtol)
        rd = np.abs(np.diag(R[:r, :r]))
        # truncation rule: every kept pivot above rtol*|R11|, next one at or below
        assert np.all(rd[1:] > rtol * rd[0])
        full = oracle.cpqr(A, 0.0)
        nxt = np.abs(full[2][r, r]) if r < min(m, n) else 0.0
        assert nxt <= rtol * rd[0] * (1 + 1e-12)
        # interpolative decomposition A ~ A[:, skel] T with T[:, skel] = I
        T = oracle.interp(R, r, piv)
        np.testing.assert_allclose(T[:, piv[:r]], np.eye(r), atol=0)
        err = np.linalg.norm(A - A[:, piv[:r]] @ T) / np.linalg.norm(A)
        assert err < 30 * rtol


def test_cpqr_brute_force_first_pivot(oracle):
    """First pivot is the column of largest norm; the first |R_kk| of a
    full pivoted QR equals the largest residual after projecting out the
    chosen columns (checked by brute force on a tiny matrix)."""
    rng = np.random.default_rng(5)
    A = crandn(rng, 6, 5)
    r, piv, R, tau = oracle.cpqr(A, 0.0)
    assert piv[0] == np.argmax(np.linalg.norm(A, axis=0))
    for k in range(1, 5):
        Q, _ = np.linalg.qr(A[:, piv[:k]])
        res = A - Q @ (Q.conj().T @ A)
        norms = np.linalg.norm(res, axis=0)
        norms[piv[:k]] = -1
        assert piv[k] == np.argmax(norms)
        assert abs(abs(R[k, k]) - norms.max()) < 1e-12 * abs(R[0, 0])


def test_judge_is_the_qr_stopping_rule_on_held_out_rows(oracle):
    """bfo_judge pinned by the pivoted QR it judges: on the sampled rows
    themselves its residual is the next pivot of a full pivoted QR of the
    sample (|R_rr|, the largest remaining column norm) and its scale the
    first (|R_00|, the largest column norm), so e_max <= eps z_max is the
    QR's own stopping rule; all columns as skeleton leave no residual, an
    empty skeleton leaves all of it, and a shifted skeleton is caught."""
    rng = np.random.default_rng(12)
    k, eps, key = 2 * np.pi, 1e-4, 987654321
    obs = np.c_[rng.uniform(0, 3, 400), rng.uniform(0, 3, 400), np.zeros(400)]
    src = np.c_[rng.uniform(0, 1, 64), rng.uniform(0, 1, 64), np.full(64, 6.0)]
    oi, si = np.arange(400), np.arange(64)
    T, skel, r, rows = oracle.compress(k, obs, oi, src, si, 30, eps, True, key)
    assert 0 < r < rows < 400
    # on the sampled rows (held-out window starting at row 0 of the prefix)
    e, z = oracle.judge(k, obs, oi, src, si, key, 0, rows, T, skel)
    sample = oracle.sample_prefix(key, 400, rows)
    _, _, R, _ = oracle.cpqr(oracle.helmholtz(k, obs, sample, src, si), 0.0)
    assert abs(e - abs(R[r, r])) <= 1e-8 * abs(R[0, 0])
    assert abs(z - abs(R[0, 0])) <= 1e-12 * abs(R[0, 0])
    assert e <= eps * z * (1 + 1e-8)
    # held-out rows: a residual of the same order, never zero
    eh, zh = oracle.judge(k, obs, oi, src, si, key, rows, 8, T, skel)
    assert 0 < eh < 1e3 * eps * zh
    # every column a skeleton column with T = I: no residual
    e0, z0 = oracle.judge(k, obs, oi, src, si, key, rows, 8, np.eye(64, dtype=complex), si)
    assert e0 == 0.0 and z0 == zh
    # empty skeleton: the whole tile is residual
    e1, z1 = oracle.judge(k, obs, oi, src, si, key, rows, 8, np.zeros((0, 64), complex), np.zeros(0, np.int32))
    assert e1 == z1 == zh
    # a planted mistake (the skeleton shifted by one source) is caught
    bad = (np.asarray(skel) + 1) % 64
    e2, _ = oracle.judge(k, obs, oi, src, si, key, rows, 8, T, bad)
    assert e2 > 100 * eps * zh


def test_judge_fallback_in_assembly(oracle, bfgpu):
    """A guess of 1 row (floor 10) with the judge: rejected compressions
    are redone by plain adaptive RRQR from max(estimator, 2 rows) rows, and
    the factorization is then within 10 tau of the dense kernel."""
    from conftest import two_squares
    obs, src = two_squares(1024, seed=6)
    bt = bfgpu.BlockTree(bfgpu.ClusterTree(obs, dims=2, leaf_size=64), bfgpu.ClusterTree(src, dims=2, leaf_size=64))
    k, tol = 2 * np.pi, 1e-3
    n = oracle.assemble(bt.testcluster, bt.trialcluster, 0, 0, k, tol).export()["comp"].shape[0]
    b = oracle.assemble(bt.testcluster, bt.trialcluster, 0, 0, k, tol, rank_mode=2, fixed=np.ones(n, np.int32),
                        judge_rows=8, judge_slack=4.0)
    ex = b.export()
    j, c = ex["judge"], ex["comp"]
    assert np.all(c[:, 3] == 1) and np.any(j == 2)
    est = oracle.rank_butterfly(tol / 3)[0]  # three levels: eps = tol / 3
    fb = c[j == 2]
    assert np.all(fb[:, 4] >= np.minimum(max(est, 20), 1024))  # at least max(estimator, 2 x 10) rows
    x = crandn(np.random.default_rng(1), 1024)
    A = oracle.helmholtz(k, obs, np.arange(1024), src, np.arange(1024))
    assert np.linalg.norm(b.matvec(x) - A @ x) / np.linalg.norm(A @ x) < 10 * tol


# ---------------------------------------------------------------- kernel
def test_helmholtz_closed_form_and_symmetry(oracle):
    k = 2 * np.pi
    x = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]])
    y = np.array([[0.25, 0.0, 0.0], [1.0, 2.0, 3.5]])
    Z = oracle.helmholtz(k, x, [0, 1], y, [0, 1])
    # quarter wavelength: exp(-i pi/2) / (4 pi / 4) = -i / pi
    assert abs(Z[0, 0] - (-1j / np.pi)) < 1e-15
    # half wavelength: exp(-i pi) / (4 pi / 2) = -1 / (2 pi)
    assert abs(Z[1, 1] - (-1 / (2 * np.pi))) < 1e-15
    # reciprocity G(x, y) = G(y, x)
    Zt = oracle.helmholtz(k, y, [0, 1], x, [0, 1])
    np.testing.assert_allclose(Z, Zt.T, rtol=1e-15)


# ---------------------------------------------------------------- estimators
def test_rank_estimators_known_answers(oracle):
    # config tolerances tau / L: 1e-4 / 4 and 1e-4 / 6 with C_tau = 4
    assert oracle.rank_butterfly(1e-4 / 4)[0] == 43
    assert oracle.rank_butterfly(1e-4 / 6)[0] == 45
    assert oracle.rank_butterfly(0.5)[0] == 10  # floor Rmin
    r, x1, x2 = oracle.rank_geometric(2 * np.pi, [0, 0, 0], [0, 0, 10.0], 1.0, 1.0, 1e-3)
    a = 2 * np.pi * 1.0 / 10.0
    assert abs(x1 - a * a) < 1e-14 and abs(x2 - np.log(1e3)) < 1e-14
    assert r == int(np.ceil(1.5 * a * a + 1.5 * np.log(1e3)))


# ---------------------------------------------------------------- row sample
def test_sample_prefix_properties(oracle):
    key = oracle.sample_key(0, 0, 0, 1, 5, 7)
    for n, m in [(10, 10), (4096, 43), (65536, 45), (100, 1)]:
        s = oracle.sample_prefix(key, n, m)
        assert len(s) == m and len(set(s.tolist())) == m
        assert s.min() >= 0 and s.max() < n
        # a longer prefix extends a shorter one (adaptive doubling)
        s2 = oracle.sample_prefix(key, n, min(2 * m, n))
        np.testing.assert_array_equal(s2[:m], s)
    # full length is a permutation
    assert sorted(oracle.sample_prefix(key, 257, 257).tolist()) == list(range(257))
    # roughly uniform first draw over many keys
    first = np.array([oracle.sample_prefix(oracle.sample_key(s, 0, 0, 0, 0, 0), 8, 1)[0]
                      for s in range(4000)])
    counts = np.bincount(first, minlength=8)
    assert counts.min() > 400 and counts.max() < 600


# ---------------------------------------------------------------- factorization
def _check_bf_against_dense(oracle, bfgpu, obs, src, tol, dims=2, leaf_o=64, leaf_s=64,
                            mode_o="balanced", mode_s="balanced"):
    To = bfgpu.ClusterTree(obs, dims=dims, leaf_size=leaf_o, mode=mode_o)
    Ts = bfgpu.ClusterTree(src, dims=dims, leaf_size=leaf_s, mode=mode_s)
    k = 2 * np.pi
    bf = oracle.assemble(To, Ts, 0, 0, k, tol)
    A = oracle.helmholtz(k, obs, np.arange(len(obs)), src, np.arange(len(src)))
    rng = np.random.default_rng(1)
    x = crandn(rng, len(src))
    y = bf.matvec(x)
    err = np.linalg.norm(y - A @ x) / np.linalg.norm(A @ x)
    return bf, A, To, Ts, err


def test_assemble_balanced_error_and_structure(oracle, bfgpu):
    obs, src = two_squares(4096)
    tol = 1e-3
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, tol)
    assert err < 10 * tol  # test_subroutines.jl:174-177
    ex = bf.export()
    tab = ex["table"]
    L = ex["n_stages"] - 1
    assert L == 4  # balanced quadtree over 4096 points, leaf 64
    counts = np.bincount(tab[:, 0], minlength=L + 1)
    # Q: 4^(L-1) leaves; every R level 4^L column slices; P: 4^(L-1)
    assert counts[0] == 4 ** (L - 1) and counts[L] == 4 ** (L - 1)
    assert all(counts[l] == 4 ** L for l in range(1, L))
    # skeletons are subsets of the candidate source points
    sk, idx = ex["skel"], ex["skel_idx"]
    for rec in sk:
        stage, s, o, ofs, ln = rec
        vals = set(Ts.cluster_values(s).tolist())
        assert set(idx[ofs:ofs + ln].tolist()) <= vals
    # column slices: R block columns equal the input skeleton lengths
    for row in tab[tab[:, 0] == 1]:
        s_in, o_in = row[4], row[3]
        ln = [r[4] for r in sk if r[0] == 0 and r[1] == s_in and r[2] == o_in]
        assert row[6] == ln[0]


def test_assemble_dense_reconstruction(oracle, bfgpu):
    obs, src = two_squares(1024, seed=4)
    tol = 1e-4
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, tol)
    D = bf.dense()
    assert np.linalg.norm(D - A) / np.linalg.norm(A) < 10 * tol


@pytest.mark.parametrize("deep_obs", [True, False])
def test_assemble_unbalanced_frozen_paths(oracle, bfgpu, deep_obs):
    """Deep observer / shallow source and the reverse exercise the source-
    and observer-frozen builders (bfassembly.jl:351-481)."""
    obs, src = two_squares(2048, seed=7)
    tol = 1e-3
    lo, ls = (32, 512) if deep_obs else (512, 32)
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, tol, leaf_o=lo, leaf_s=ls)
    assert len(To.cluster_levels()) != len(Ts.cluster_levels())
    assert err < 10 * tol


def test_assemble_maxpoints_tree_pads_leaves(oracle, bfgpu):
    """Unbalanced (max-points) trees: shallow leaves are padded down
    (traverseandpad, auxillaries.jl:340-350) and give identity blocks."""
    rng = np.random.default_rng(11)
    n = 2000
    nd = int(0.85 * n)
    # a sparse observer patch beside the dense square -> shallow leaves
    obs = np.c_[np.r_[rng.uniform(0, 16, nd), rng.uniform(16, 24, n - nd)],
                np.r_[rng.uniform(0, 16, nd), rng.uniform(0, 8, n - nd)], np.zeros(n)]
    src = np.c_[rng.uniform(0, 16, n), rng.uniform(0, 16, n), np.full(n, 32.0)]
    tol = 1e-4
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, tol, leaf_o=32,
                                                 leaf_s=128, mode_o="maxpoints", mode_s="maxpoints")
    ex = bf.export()
    assert ex["table"][:, 7].sum() > 0  # identity blocks present
    assert err < 10 * tol


def test_adjoint_transpose_and_multirhs(oracle, bfgpu):
    obs, src = two_squares(1024, seed=2)
    tol = 1e-4
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, tol)
    rng = np.random.default_rng(9)
    x = crandn(rng, 1024)
    bh, bt = bf.transform(1), bf.transform(0)
    ya = bh.matvec(x)
    yt = bt.matvec(x)
    assert np.linalg.norm(ya - A.conj().T @ x) / np.linalg.norm(A.conj().T @ x) < 1e-2
    assert np.linalg.norm(yt - A.T @ x) / np.linalg.norm(A.T @ x) < 1e-2
    # exact algebra of the factorization: <B x, y> = <x, B^H y>, B^T = conj(B^H conj)
    xx, yy = crandn(rng, 1024), crandn(rng, 1024)
    lhs = np.vdot(yy, bf.matvec(xx))
    rhs = np.vdot(bh.matvec(yy), xx)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    np.testing.assert_allclose(bt.matvec(xx), np.conj(bh.matvec(np.conj(xx))), rtol=1e-13, atol=1e-15)
    # multi-RHS equals column by column; alpha, beta semantics
    X = crandn(rng, 1024, 3)
    Y = bf.matvec(X)
    for j in range(3):
        np.testing.assert_allclose(Y[:, j], bf.matvec(X[:, j]), rtol=0, atol=0)
    y0 = crandn(rng, 1024)
    y1 = bf.matvec(x, alpha=2 - 1j, beta=0.5j, y=y0)
    np.testing.assert_allclose(y1, (2 - 1j) * bf.matvec(x) + 0.5j * y0, rtol=1e-13, atol=1e-14)


def test_planted_mistake_is_caught(oracle, bfgpu):
    """A transposed operand planted in the oracle's dense product is caught
    by the dense-comparison bar the tests use."""
    obs, src = two_squares(1024, seed=2)
    bf, A, To, Ts, err = _check_bf_against_dense(oracle, bfgpu, obs, src, 1e-3)
    rng = np.random.default_rng(0)
    x = crandn(rng, 1024)
    bad = np.linalg.norm(bf.matvec(x) - A.T @ x) / np.linalg.norm(A @ x)
    assert bad > 1e-2


# ---------------------------------------------------------------- far field
def test_nearandfar_tiles_the_matrix(oracle, bfgpu):
    from conftest import fibonacci_sphere
    pts = fibonacci_sphere(3000, 2.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    for kind, param in [(0, 1.0), (1, 2.0)]:
        far, near = oracle.nearandfar(T, T, kind, param)
        cover = np.zeros((3000, 3000), np.int8)
        for o, s in np.r_[far, near]:
            cover[np.ix_(T.cluster_values(o), T.cluster_values(s))] += 1
        assert cover.min() == 1 and cover.max() == 1
        # admissibility closed form on every far pair
        for o, s in far:
            d = np.linalg.norm(T.center[o] - T.center[s])
            ro, rs = T.halfsize[o] * np.sqrt(3), T.halfsize[s] * np.sqrt(3)
            if kind == 1:
                assert d > param * (ro + rs)
            else:
                gap = np.maximum(np.abs(T.center[o] - T.center[s]) - (ro + rs), 0)
                assert np.linalg.norm(gap) > param * max(ro, rs)


def test_near_dense_places_every_block(oracle, bfgpu):
    """The near-field oracle: coincident points take selfterm; near blocks
    plus the far blocks written densely rebuild the whole kernel matrix
    (each entry once); a symmetric pair list gives a symmetric matrix
    (reciprocity), so a transposed placement fails."""
    from conftest import fibonacci_sphere
    k = 2 * np.pi
    x = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.0]])
    y = np.array([[0.0, 0.0, 0.0], [1.0, 2.0, 3.25]])
    Z = oracle.near_block(k, x, [0, 1], y, [0, 1], selfterm=0.5 + 2j)
    assert Z[0, 0] == 0.5 + 2j
    assert abs(Z[1, 1] - (-1j / np.pi)) < 1e-15  # quarter wavelength
    pts = fibonacci_sphere(1500, 2.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    far, near = oracle.nearandfar(T, T, 1, 2.0)
    N = oracle.near_dense(k, T, T, near, selfterm=0.3)
    F = oracle.near_dense(k, T, T, far)
    idx = np.arange(1500, dtype=np.int32)
    G = oracle.near_block(k, pts, idx, pts, idx, selfterm=0.3)
    np.testing.assert_allclose(N + F, G, rtol=1e-15, atol=0)
    np.testing.assert_allclose(N, N.T, rtol=1e-15, atol=0)
    assert np.count_nonzero(F) > 0 and np.count_nonzero(N) + np.count_nonzero(F) == N.size
    assert np.all(np.diag(N) == 0.3)
