"""Streaming forward kernel (variant 9) against the oracle.

At one right-hand side a forward level can run the streaming kernel: runs
of consecutive single-block tasks whose blocks lie back to back in the
factor buffer are copied, 4 or 8 KiB at a time, into the wavefront's LDS
and computed there; tasks it does not take (identity blocks, several
contributions, blocks larger than a chunk, rows of y) run the dense kernel
in the same level.  Bars (written in each test): FP64 products equal the
oracle's matvec of the same factorization to 1e-12 relative; FP32 products
equal the oracle's FP64 product of the same FP32 factors to 2e-6 (FP32
rounding of the level chain; the kernels differ in summation order only).
"""
import sys

import numpy as np
import pytest

from conftest import crandn, fibonacci_sphere, two_squares

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


def _force_stream_tr(bfgpu, g, kib, skip_last=False):
    """Every transposed table (and partial-sum table) on the streaming kernel
    over the transposed copy; skip_last leaves the last level (matrix-free,
    not copied) alone."""
    sh = bfgpu.ButterflyFactorization.tile_shape(4, 0, 0, 7, 0, 0, 1 if kib == 4 else 0)
    v = g.tuning().copy()
    n = len(v) // 8
    for L in range(n - 1 if skip_last else n):
        for q in (2, 3):
            if v[8 * L + 2 * q] >= 0:
                v[8 * L + 2 * q] = 9
                v[8 * L + 2 * q + 1] = sh
    g.set_tuning(v)


def _check(g, refs, X, bar):
    """One RHS on every direction (the streaming kernel runs forward at one
    RHS), and 2 / 8 RHS forward (the dense kernel takes those)."""
    for name, want in refs.items():
        op = {"N": g, "C": g.adjoint(), "T": g.transpose()}[name]
        got = op @ X[:, 0].copy()
        assert rel(got.astype(np.complex128), want[:, 0]) < bar, (name, 1)
    for nr in (2, 8):
        got = g @ np.asfortranarray(X[:, :nr])
        assert rel(got.astype(np.complex128), refs["N"][:, :nr]) < bar, ("N", nr)


def test_stream_two_squares_fp64_matches_oracle(bfgpu, oracle, ctx):
    say("two squares N=4096 fp64: streaming kernel, 8 and 4 KiB chunks")
    n = 4096
    obs, src = two_squares(n, seed=2)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(To, Ts), 0, 0, K, 1e-4, ctx=ctx)
    imp = oracle.import_bf(To, Ts, g.export())
    adj, tra = imp.transform(1), imp.transform(0)

    def cols(f):
        return lambda X: np.stack([f(X[:, j]) for j in range(X.shape[1])], axis=1)

    X = crandn(np.random.default_rng(21), n, 8)
    refs = {"N": cols(imp.matvec)(X), "C": cols(adj.matvec)(X), "T": cols(tra.matvec)(X)}
    assert g.get_option("stream") == 0
    for kib in (8, 4):
        g.set_option("stream", kib)
        assert g.get_option("stream") == kib
        st = g.stream_stats()
        say("stream stats", st.tolist())
        # every forward level built with this chunk size (blocks larger
        # than a chunk are left to the dense kernel)
        assert (st[:, 0] == kib).all()
        # the last level writes rows of y: none of its tasks stream
        assert st[-1, 2] == 0 and st[-1, 3] > 0
        _check(g, refs, X, 1e-12)
        # alpha / beta on the forward product
        Y0 = crandn(np.random.default_rng(22), n, 1)[:, 0]
        a, b = 0.5 - 0.25j, -1.5 + 0.5j
        got = g.mul(Y0.copy(), X[:, 0].copy(), a, b)
        assert rel(got, a * refs["N"][:, 0] + b * Y0) < 1e-12
    # the tuning round trip keeps variant 9 and the chunk size
    v = g.tuning().copy()
    assert v[8 * 1] == 9
    g.set_option("stream", 0)
    assert g.get_option("stream") == 0 and all(g.tuning()[8 * L] == 3 for L in range(len(g)))
    _check(g, refs, X, 1e-12)
    g.set_tuning(v)
    assert g.get_option("stream") == 4
    _check(g, refs, X, 1e-12)
    # graph replay of the streaming launches
    g.set_option("graph", 1)
    for _ in range(3):
        got = g @ X[:, 0].copy()
        assert rel(got, refs["N"][:, 0]) < 1e-12
    g.set_option("graph", 0)
    # the autotuner's choice (it times the streaming kernel at one RHS)
    g.autotune(nrhs=1, iters=2)
    _check(g, refs, X, 1e-12)
    with pytest.raises(bfgpu.ArgumentError):
        g.set_option("stream", 6)
    # adjoint and transpose on the streaming kernel over the transposed copy
    # (an item's blocks back to back, its several input vectors copied one
    # after the other); several RHS take the dense kernel over the copy
    g.set_option("tcopy", 1)
    for kib in (8, 4):
        say("transposed tables streaming over the copy,", kib, "KiB")
        _force_stream_tr(bfgpu, g, kib)
        assert all(g.tuning()[8 * L + 4] == 9 for L in range(len(g)))
        for name in ("C", "T"):
            op = g.adjoint() if name == "C" else g.transpose()
            got = op @ X[:, 0].copy()
            assert rel(got, refs[name][:, 0]) < 1e-12, (name, kib)
            got = op @ np.asfortranarray(X[:, :3])
            assert rel(got, refs[name][:, :3]) < 1e-12, (name, kib, 3)
        Y0 = crandn(np.random.default_rng(24), n, 1)[:, 0]
        a, b = 0.5 - 0.25j, -1.5 + 0.5j
        got = g.adjoint().mul(Y0.copy(), X[:, 0].copy(), a, b)
        assert rel(got, a * refs["C"][:, 0] + b * Y0) < 1e-12
    g.set_option("tcopy", 0)  # released: variant 9 of the transposed tables goes back to the lazy kernels
    assert all(g.tuning()[8 * L + 4] != 9 for L in range(len(g)))
    assert rel(g.adjoint() @ X[:, 0].copy(), refs["C"][:, 0]) < 1e-12


def test_stream_sphere_far_field_matches_oracle(bfgpu, oracle, ctx):
    say("far field sphere N=6000: streaming kernel, fp64 and fp32 (matrix-free last level)")
    n = 6000
    pts = fibonacci_sphere(n, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    X = crandn(np.random.default_rng(23), n, 8)
    for acc, bar in ((np.complex128, 1e-12), (np.complex64, 2e-6)):
        op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4,
                                    admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=acc,
                                    farfieldonly=True, ctx=ctx)
        imps = [oracle.import_bf(T, T, op.export(i)) for i in range(op.n_factorizations())]
        adj = [b.transform(1) for b in imps]
        tra = [b.transform(0) for b in imps]

        def cols(bfs):
            return lambda Z: np.stack([oracle.matvec_many(bfs, Z[:, j], nthreads=8, rows=n)
                                       for j in range(Z.shape[1])], axis=1)

        Xa = X.astype(acc)
        Xr = Xa.astype(np.complex128)
        refs = {"N": cols(imps)(Xr), "C": cols(adj)(Xr[:, :1]), "T": cols(tra)(Xr[:, :1])}
        dense = op @ Xa[:, 0].copy()
        for kib in (8, 4):
            say(np.dtype(acc).name, "chunks", kib, "KiB")
            op.set_option("stream", kib)
            st = op.stream_stats()
            say("stream stats", st.tolist())
            # the levels below the last hold small single-block tasks: most stream
            assert st[:-1, 2].sum() > st[:-1, 3].sum()
            _check(op, refs, Xa, bar)
            got = op @ Xa[:, 0].copy()
            # the same FP32 products in another summation order
            assert rel(got.astype(np.complex128), dense.astype(np.complex128)) < bar
        op.autotune(nrhs=1, iters=2)
        _check(op, refs, Xa, bar)
        # the adjoint / transpose streaming over the transposed copy (the
        # FP32 last level stays matrix-free and is not copied)
        op.set_option("tcopy", 1)
        _force_stream_tr(bfgpu, op, 8, skip_last=acc == np.complex64)
        for name in ("C", "T"):
            view = op.adjoint() if name == "C" else op.transpose()
            got = view @ Xa[:, 0].copy()
            assert rel(got.astype(np.complex128), refs[name][:, 0]) < bar, name
