"""GPU parity of the near field (petrovgalerkinbf.jl:55-119, the near mul!
of matrixvector.jl:31-37) and of the full operator near + far.

Bars (written in each test): FP64 near blocks equal the oracle's blocks to
1e-13 relative and FP64 near products the oracle's dense near matrix times
x to 1e-12; FP32 to 2e-6 (single-precision entries, sums of a few thousand
terms); the full FP64 operator matches the dense kernel matrix to the
compression tolerance.
"""
import sys

import numpy as np
import pytest

from conftest import crandn, fibonacci_sphere

pytestmark = pytest.mark.gpu

K = 2 * np.pi


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ctx(bfgpu):
    return bfgpu.default_context()


@pytest.fixture(scope="module")
def sphere(bfgpu):
    pts = fibonacci_sphere(3000, 3.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    far, near = bfgpu.nearandfar(bt, bfgpu.CenterDistanceAdmissibility(2.0))
    return T, bt, far, near


def test_near_blocks_match_oracle(bfgpu, oracle, ctx, sphere):
    T, bt, far, near = sphere
    say("near blocks sphere N=3000 fp64/fp32, pairs", near.shape[0])
    N64 = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, near, ctx=ctx)
    N32 = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, near, acctype=np.complex64, ctx=ctx)
    info = N64.info()
    assert info["pairs"] == near.shape[0] and info["rows"] == info["cols"] == 3000
    assert info["direct"] == 1  # leaf rows are disjoint and cover every point
    assert info["entries"] == sum(len(T.cluster_values(o)) * len(T.cluster_values(s)) for o, s in near)
    n_self = 0
    for i, (o, s) in enumerate(near):
        want = oracle.near_block(K, T.points, T.cluster_values(o), T.points, T.cluster_values(s))
        n_self += int(np.sum(want == 0))
        got = N64.block(i)
        assert got.shape == want.shape
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want), (i, o, s)
        got32 = N32.block(i).astype(np.complex128)
        assert np.linalg.norm(got32 - want) <= 2e-6 * np.linalg.norm(want), (i, o, s)
    assert n_self == 3000  # one coincident point per row: the diagonal


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_near_products_match_oracle(bfgpu, oracle, ctx, sphere, prec):
    T, bt, far, near = sphere
    dt = np.complex128 if prec == "fp64" else np.complex64
    bar = 1e-12 if prec == "fp64" else 2e-6
    say("near products sphere N=3000", prec)
    Nf = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, near, acctype=dt, selfterm=0.5 - 0.25j,
                               ctx=ctx)
    D = oracle.near_dense(K, T, T, near, selfterm=0.5 - 0.25j)
    rng = np.random.default_rng(3)
    a, b = 0.7 - 0.2j, -0.3 + 1.1j
    # adjoint / transpose through the transposed copy, then the two-pass
    # product over the stored blocks
    # complex FP32 from 8 RHS runs on the matrix cores (passes of <= 64 RHS)
    for copy in (True, False):
        Nf.set_transposed_copy(copy)
        for nrhs in (1, 2, 3, 9, 40, 70):
            X = np.asfortranarray(crandn(rng, 3000, nrhs).astype(dt))
            Y0 = np.asfortranarray(crandn(rng, 3000, nrhs).astype(dt))
            Xd = X.astype(np.complex128)
            for name, op, M in (("N", Nf, D), ("T", Nf.transpose(), D.T), ("C", Nf.adjoint(), D.conj().T)):
                say("  copy", copy, "nrhs", nrhs, "op", name)
                y = op.mul(Y0.copy(order="F"), X, a, b)
                want = a * (M @ Xd) + b * Y0.astype(np.complex128)
                assert rel(y, want) < bar, (copy, nrhs, name)
                y0 = op @ X
                assert rel(y0, M @ Xd) < bar, (copy, nrhs, name)
    Nf.set_transposed_copy(True)
    # the vector-ALU path for many RHS agrees with the matrix cores
    X = np.asfortranarray(crandn(rng, 3000, 24).astype(dt))
    on = Nf @ X
    Nf.set_mfma(0)
    off = Nf @ X
    Nf.set_mfma(8)
    assert rel(on, off) < bar
    # one vector, beta = 0 ignores what y held (NaN included)
    x = crandn(rng, 3000).astype(dt)
    y = np.full(3000, np.nan, dt)
    Nf.mul(y, x, 1.0, 0.0)
    assert rel(y, D @ x.astype(np.complex128)) < bar


def test_near_overlapping_pairs_two_pass(bfgpu, oracle, ctx, sphere):
    """Pairs whose observer rows overlap (a node and its child), repeated
    pairs and pairs of unequal trees take the two-pass product; the oracle
    adds overlapping blocks, as the BlockSparseMatrix product does."""
    T, bt, far, near = sphere
    leaves = [v for v in range(T.n_nodes) if T.cluster_isleaf(v)]
    child = leaves[5]
    parent = T.cluster_parent(child)
    pairs = np.array([[child, leaves[0]], [parent, leaves[1]], [child, leaves[0]], [leaves[7], parent],
                      [T.cluster_root(), leaves[2]]], np.int32)
    say("near overlapping pairs", pairs.tolist())
    for dt, bar in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        Nf = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, pairs, acctype=dt, ctx=ctx)
        assert Nf.info()["direct"] == 0
        D = oracle.near_dense(K, T, T, pairs)
        rng = np.random.default_rng(4)
        # 9 RHS: the matrix cores take only direct fields, so this is the
        # vector two-pass product even in FP32
        X = np.asfortranarray(crandn(rng, 3000, 9).astype(dt))
        Y0 = np.asfortranarray(crandn(rng, 3000, 9).astype(dt))
        for copy in (True, False):
            Nf.set_transposed_copy(copy)
            for op, M in ((Nf, D), (Nf.adjoint(), D.conj().T), (Nf.transpose(), D.T)):
                y = op.mul(Y0.copy(order="F"), X, 2.0, -1.0)
                assert rel(y, 2.0 * (M @ X.astype(np.complex128)) - Y0) < bar, copy

    say("near field of two different trees")
    rng = np.random.default_rng(5)
    src = rng.uniform(-1, 1, (700, 3))
    obs = rng.uniform(-1, 1, (500, 3)) + np.array([0.5, 0.0, 0.0])
    To = bfgpu.ClusterTree(obs, dims=3, leaf_size=32, mode="maxpoints")
    Ts = bfgpu.ClusterTree(src, dims=3, leaf_size=32, mode="maxpoints")
    b2 = bfgpu.BlockTree(To, Ts)
    f2, n2 = bfgpu.nearandfar(b2, bfgpu.CenterDistanceAdmissibility(2.0))
    assert n2.shape[0] > 0
    Nf = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), b2, n2, ctx=ctx)
    assert Nf.shape == (500, 700)
    D = oracle.near_dense(K, To, Ts, n2)
    x = crandn(rng, 700)
    assert rel(Nf @ x, D @ x) < 1e-12
    z = crandn(rng, 500)
    assert rel(Nf.adjoint() @ z, D.conj().T @ z) < 1e-12


def test_near_empty_and_device_tensors(bfgpu, oracle, ctx, sphere):
    import torch
    T, bt, far, near = sphere
    say("near field with no pairs: y = beta y")
    E = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, np.zeros((0, 2), np.int32), ctx=ctx)
    rng = np.random.default_rng(6)
    y0 = crandn(rng, 3000)
    y = E.mul(y0.copy(), crandn(rng, 3000), 1.0, 0.5j)
    assert np.array_equal(y, 0.5j * y0)
    assert np.array_equal(E.adjoint().mul(y0.copy(), crandn(rng, 3000), 1.0, 0.5j), 0.5j * y0)

    say("near products on torch device tensors, fp32, 1 and 6 RHS")
    Nf = bfgpu.BlockSparseNear(bfgpu.HelmholtzKernelMatrix(K), bt, near, acctype=np.complex64, ctx=ctx)
    D = oracle.near_dense(K, T, T, near)
    for nrhs in (1, 6):
        X = crandn(rng, 3000, nrhs).astype(np.complex64)
        xt = torch.from_numpy(np.ascontiguousarray(X.T)).cuda().t()  # column-major view
        if nrhs == 1:
            xt = xt[:, 0].contiguous()
        yt = Nf @ xt
        torch.cuda.synchronize()
        got = yt.cpu().numpy().reshape(3000, -1) if nrhs > 1 else yt.cpu().numpy()
        want = D @ X.astype(np.complex128)
        assert rel(got, want if nrhs > 1 else want[:, 0]) < 2e-6


def test_full_operator_matches_dense_matrix(bfgpu, oracle, ctx):
    """PetrovGalerkinBF with its near field applies the whole kernel matrix:
    near + far against the dense matrix (coincident points = selfterm)."""
    pts = fibonacci_sphere(4096, 3.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    say("full operator sphere N=4096 fp64 tol 1e-6")
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-6,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), selfterm=0.25, ctx=ctx)
    assert op.nearinteractions is not None and op.farinteractions.shape[0] > 0
    idx = np.arange(4096, dtype=np.int32)
    G = oracle.near_block(K, pts, idx, pts, idx, selfterm=0.25)
    rng = np.random.default_rng(7)
    X = np.asfortranarray(crandn(rng, 4096, 3))
    Y = op @ X
    assert rel(Y, G @ X) < 1e-5
    assert rel(op.adjoint() @ X, G.conj().T @ X) < 1e-5
    assert rel(op.transpose() @ X[:, 0], G.T @ X[:, 0]) < 1e-5
    # the far field alone misses the near blocks
    assert rel(op.farfield() @ X, G @ X) > 1e-2
    # the adjoint constructed explicitly applies as the adjoint view
    Ah = op.adjoint().materialize()
    assert rel(Ah @ X, op.adjoint() @ X) < 1e-12
    assert rel(Ah.adjoint() @ X, Y) < 1e-12
    # alpha / beta through both parts
    Y0 = np.asfortranarray(crandn(rng, 4096, 3))
    got = op.mul(Y0.copy(order="F"), X, 0.5, 2.0 - 1.0j)
    assert rel(got, 0.5 * (G @ X) + (2.0 - 1.0j) * Y0) < 1e-5

    say("FP32 full operator sphere N=4096 tol 1e-4")
    op32 = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4, acctype=np.complex64,
                                  admissibility=bfgpu.CenterDistanceAdmissibility(2.0), selfterm=0.25, ctx=ctx)
    y32 = op32 @ X[:, 0].astype(np.complex64)
    assert rel(y32, G @ X[:, 0]) < 1e-3


def test_near_only_operator(bfgpu, oracle, ctx):
    """A problem too small for any admissible block is all near field."""
    pts = fibonacci_sphere(200, 1.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    far, near = bfgpu.nearandfar(bt, bfgpu.CenterDistanceAdmissibility(2.0))
    say("near-only operator N=200, far pairs", far.shape[0])
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), ctx=ctx)
    idx = np.arange(200, dtype=np.int32)
    G = oracle.near_block(K, pts, idx, pts, idx)
    x = crandn(np.random.default_rng(8), 200)
    if far.shape[0] == 0:
        assert op.shape == (200, 200)
    assert rel(op @ x, G @ x) < 1e-4


def _power_norm(apply, apply_h, n, rng, tol=1e-4, itmax=100):
    """sqrt of the largest eigenvalue of M^H M by power iteration (the
    spectral-norm estimate of test_pgalerkin.jl)."""
    v = crandn(rng, n)
    v /= np.linalg.norm(v)
    s_old, s_new = 1.0, 1.0
    for i in range(itmax):
        if i >= 3 and abs(np.sqrt(s_old) - np.sqrt(s_new)) / np.sqrt(s_old) <= tol:
            break
        s_old = s_new
        x = apply_h(apply(v))
        s_new = np.linalg.norm(x)
        if s_new < 1e-14:
            return 0.0
        v = x / s_new
    return np.sqrt(s_new)


def test_operator_spectral_error_within_ten_tolerances(bfgpu, oracle, ctx):
    """The global operator's check of test_pgalerkin.jl: relative spectral
    difference ||H - A||_2 / ||A||_2 of the whole operator (near + far) to
    the dense matrix, by power iteration on products with H and H', below
    10 tol at tol = 1e-3."""
    pts = fibonacci_sphere(4096, 2.5)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    say("spectral error of the whole operator, sphere N=4096, tol 1e-3, fp64 and fp32")
    idx = np.arange(4096, dtype=np.int32)
    A = oracle.near_block(K, pts, idx, pts, idx)
    Ah = A.conj().T.copy()
    rng = np.random.default_rng(21)
    nA = _power_norm(lambda v: A @ v, lambda w: Ah @ w, 4096, rng)
    for dt in (np.complex128, np.complex64):
        H = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(T, T), K, tol=1e-3,
                                   admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=dt, ctx=ctx)
        Hh = H.adjoint()
        cast = lambda v: v.astype(dt)
        err = _power_norm(lambda v: (H @ cast(v)).astype(np.complex128) - A @ v,
                          lambda w: (Hh @ cast(w)).astype(np.complex128) - Ah @ w, 4096, rng) / nA
        say("  ", np.dtype(dt).name, "relative spectral error", err)
        assert err < 10 * 1e-3
