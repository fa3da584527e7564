"""Kernel variants of the matvec levels against the oracle.

Every launch table can run on the workgroup-per-task kernel (variant 0), the
wavefront-per-task kernel (variant 1, forward), the tiled kernel (variant 2)
or the dense kernel (variant 3: forward packed rows; transposed a lane per
output column) in any legal shape, and the transposed tables on the forward
dense kernel over a transposed copy of the blocks (variant 8); observer-leaf sums over many
factorizations are split into partial sums plus a combine pass.  Bar
(written in each test): every variant and shape gives the oracle's FP64
matvec of the same factorization to 1e-12 relative, in all three
directions and for 1, 2, 3 and 8 right-hand sides.
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


def _shapes(bfgpu):
    """A spread of legal tiled shapes (rows x groups, both group modes)."""
    out = []
    for rp in range(0, 7):
        for gs in sorted({0, min(2, 6 - rp), 6 - rp}):
            for cs in (0, 1):
                if (cs and gs == 0) or rp + gs < 2:
                    continue
                out.append(bfgpu.ButterflyFactorization.tile_shape(rp, gs, cs, 8))
    out.append(bfgpu.ButterflyFactorization.tile_shape(3, 1, 0, 1))  # one workgroup per CU
    # contiguous task range per XCD (small levels launch fewer than 8
    # workgroups and fall back to the plain mapping)
    for rp, gs, cs, wg in ((3, 1, 0, 8), (2, 2, 1, 8), (4, 0, 0, 1), (6, 0, 0, 2)):
        out.append(bfgpu.ButterflyFactorization.tile_shape(rp, gs, cs, wg, 1))
    return out


def _dense_shapes(bfgpu):
    """Dense shapes: lanes per task 4..64, both group modes (transposed
    lane-per-column kernel: inputs by permutes or from LDS), 4 or 8 loads
    per pass, both task-to-XCD mappings."""
    sh = bfgpu.ButterflyFactorization.tile_shape
    out = [sh(lp, 0, cs, 8, 0, q4) for lp in range(2, 7) for cs in (0, 1) for q4 in (0, 1)]
    out += [sh(4, 0, 0, 8, 1, 0), sh(5, 0, 1, 2, 1, 1), sh(6, 0, 0, 1, 1, 0)]
    return out


def _force_tcopy(g, sh, skip_last=False):
    """Every transposed table (and partial-sum table) on variant 8, the
    forward dense kernel over the transposed copy, with dense shape sh;
    skip_last leaves the last level (matrix-free, not copied) alone."""
    v = g.tuning().copy()
    n = len(v) // 8
    for L in range(n - 1 if skip_last else n):
        for q in (2, 3):
            if v[8 * L + 2 * q] >= 0:
                v[8 * L + 2 * q] = 8
                v[8 * L + 2 * q + 1] = sh
    g.set_tuning(v)


def _check_all_directions(g, refs, X):
    """refs: op name -> the oracle's product of X (columns independent)."""
    for name, want in refs.items():
        op = {"N": g, "C": g.adjoint(), "T": g.transpose()}[name]
        for nr in (1, 2, 3, 8):
            Xs = np.asfortranarray(X[:, :nr])
            got = op @ Xs
            if nr == 1 and got.ndim == 1:
                got = got[:, None]
            assert rel(got, want[:, :nr]) < 1e-12, (name, nr)


def test_variants_two_squares_match_oracle(bfgpu, oracle, ctx):
    say("two squares N=4096 fp64: kernel variants and tiled shapes")
    n = 4096
    obs, src = two_squares(n, seed=2)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    bt = bfgpu.BlockTree(To, Ts)
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bt, 0, 0, K, 1e-4, ctx=ctx)
    imp = oracle.import_bf(To, Ts, g.export())
    adj, tra = imp.transform(1), imp.transform(0)

    def cols(f):
        return lambda X: np.stack([f(X[:, j]) for j in range(X.shape[1])], axis=1)

    X = crandn(np.random.default_rng(7), n, 8)
    refs = {"N": cols(imp.matvec)(X), "C": cols(adj.matvec)(X), "T": cols(tra.matvec)(X)}
    # traffic columns: the first level reads every source value once, the
    # last one reads and writes every row of y
    st = g.launch_stats()
    assert st[0, 8] == n and st[-1, 9] == 2 * n
    assert (st[:, 3] > 0).all() and st[:, 3].sum() * 16 == g.info().factor_bytes
    for v in (0, 1, 2, 3):
        say("variant", v)
        g.set_option("variant", v)
        assert g.get_option("variant") == v
        _check_all_directions(g, refs, X)
    for sh in _shapes(bfgpu):
        say("tile", hex(sh))
        g.set_option("tile", sh)
        assert g.get_option("tile") == sh and g.get_option("variant") == 2
        _check_all_directions(g, refs, X)
    for sh in _dense_shapes(bfgpu):
        say("dense", hex(sh))
        g.set_option("dense", sh)
        assert g.get_option("dense") == sh
        _check_all_directions(g, refs, X)
    # adjoint / transpose through the transposed copy of the blocks
    g.set_option("tcopy", 1)
    assert g.get_option("tcopy") == 1
    for sh in _dense_shapes(bfgpu)[::3]:
        say("transposed copy, dense", hex(sh))
        _force_tcopy(g, sh)
        assert all(g.tuning()[8 * L + 4] == 8 for L in range(len(g)))
        _check_all_directions(g, refs, X)
    g.set_option("tcopy", 0)  # released: the transposed tables go back to the lazy kernels
    assert g.get_option("tcopy") == 0 and all(g.tuning()[8 * L + 4] != 8 for L in range(len(g)))
    _check_all_directions(g, refs, X)
    with pytest.raises(bfgpu.ArgumentError):  # no copy to run on
        _force_tcopy(g, bfgpu.ButterflyFactorization.tile_shape(4, 0))
    g.set_option("tcopy", 1)
    # the autotuner's choice (both directions) keeps the product
    ms, var = g.autotune(nrhs=1, iters=2)
    assert len(ms) == len(var) == len(g)
    _check_all_directions(g, refs, X)
    with pytest.raises(bfgpu.ArgumentError):
        g.set_option("tile", bfgpu.ButterflyFactorization.tile_shape(4, 3))
    with pytest.raises(bfgpu.ArgumentError):  # fewer than 4 lanes per task
        g.set_option("dense", bfgpu.ButterflyFactorization.tile_shape(1, 0))
    with pytest.raises(bfgpu.ArgumentError):
        g.set_option("variant", 4)


def test_variants_sphere_partial_sums_match_oracle(bfgpu, oracle, ctx):
    say("far field sphere N=6000 fp64: partial sums and kernel variants")
    n = 6000
    pts = fibonacci_sphere(n, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), farfieldonly=True,
                                ctx=ctx)
    fw = op.launch_stats()
    bw = op.adjoint().launch_stats()
    say("far pairs", len(op.farinteractions), "partial items fwd", fw[-1, 7], "adj", bw[-1, 7])
    # observer-leaf sums of the last forward level and source-leaf sums of
    # the last adjoint level are split into partial sums at this size
    assert fw[-1, 7] > 0 and bw[-1, 7] > 0
    say("importing", len(op.farinteractions), "factorizations into the oracle")
    imps = [oracle.import_bf(T, T, op.export(i)) for i in range(len(op.farinteractions))]
    adj = [b.transform(1) for b in imps]
    tra = [b.transform(0) for b in imps]

    def cols(bfs):
        return lambda X: np.stack([oracle.matvec_many(bfs, X[:, j], nthreads=8, rows=n)
                                   for j in range(X.shape[1])], axis=1)

    X = crandn(np.random.default_rng(8), n, 8)
    refs = {"N": cols(imps)(X), "C": cols(adj)(X), "T": cols(tra)(X)}
    for v in (0, 1, 2, 3):
        say("variant", v)
        op.set_option("variant", v)
        _check_all_directions(op, refs, X)
    for sh in _shapes(bfgpu)[::3] + _shapes(bfgpu)[-4:]:
        say("tile", hex(sh))
        op.set_option("tile", sh)
        _check_all_directions(op, refs, X)
    for sh in _dense_shapes(bfgpu)[::2]:
        say("dense", hex(sh))
        op.set_option("dense", sh)
        _check_all_directions(op, refs, X)
    for sh in _dense_shapes(bfgpu)[1::5]:  # partial-sum tables over the transposed copy too
        say("transposed copy, dense", hex(sh))
        _force_tcopy(op, sh)
        _check_all_directions(op, refs, X)
    op.autotune(nrhs=1, iters=2)
    _check_all_directions(op, refs, X)


def test_rhs_across_lanes_matches_oracle(bfgpu, oracle, ctx):
    """Forward matvecs with many right-hand sides put them across lanes
    (64 per sweep, factor entries read once per sweep): 16, 17, 64 and 70
    columns (two sweeps) against the oracle, single block and far field
    (partial sums), alpha/beta included; switching the path off gives the
    same product."""
    say("two squares N=4096 fp64: right-hand sides across lanes")
    n = 4096
    obs, src = two_squares(n, seed=4)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(To, Ts), 0, 0, K, 1e-4, ctx=ctx)
    imp = oracle.import_bf(To, Ts, g.export())
    rng = np.random.default_rng(11)
    X = np.asfortranarray(crandn(rng, n, 70))
    want = np.stack([imp.matvec(X[:, j]) for j in range(70)], axis=1)
    assert g.get_option("rhs_lanes") == 32
    g.set_option("rhs_lanes", 16)
    for nr in (16, 17, 64, 70):
        got = g @ np.asfortranarray(X[:, :nr])
        assert rel(got, want[:, :nr]) < 1e-12, nr
    Y0 = np.asfortranarray(crandn(rng, n, 17))
    a, b = 0.5 - 0.25j, -1.5 + 0.5j
    got = g.mul(Y0.copy(order="F"), np.asfortranarray(X[:, :17]), a, b)
    assert rel(got, a * want[:, :17] + b * Y0) < 1e-12
    g.set_option("rhs_lanes", 0)
    off = g @ np.asfortranarray(X[:, :17])
    g.set_option("rhs_lanes", 16)
    assert rel(off, want[:, :17]) < 1e-12

    say("far field sphere N=6000 fp64: right-hand sides across lanes, partial sums")
    pts = fibonacci_sphere(6000, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(T, T), K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), farfieldonly=True,
                                ctx=ctx)
    assert op.launch_stats()[-1, 7] > 0
    imps = [oracle.import_bf(T, T, op.export(i)) for i in range(len(op.farinteractions))]
    Xs = np.asfortranarray(crandn(rng, 6000, 20))
    want = np.stack([oracle.matvec_many(imps, Xs[:, j], nthreads=8, rows=6000) for j in range(20)], axis=1)
    op.set_option("rhs_lanes", 16)
    assert rel(op @ Xs, want) < 1e-12

    say("far field sphere N=6000 fp32, 40 RHS: skeleton vectors capped below 16 interleaved columns")
    op32 = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(T, T), K, tol=1e-4,
                                  admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                  farfieldonly=True, ctx=ctx)
    op32.set_option("pfree", 0)  # stored last-level blocks on both paths: FP32 arithmetic compared at 2e-6
    X40 = np.asfortranarray(crandn(rng, 6000, 40).astype(np.complex64))
    lanes = op32 @ X40                                  # right-hand sides across lanes (matrix cores)
    col_kib = op32.info().slot_len * 8 / 1024
    op32.set_option("slot_cap", int(5.5 * col_kib))    # 5 columns fit, 16 interleaved do not
    assert op32.get_option("slot_cap") == int(5.5 * col_kib)
    split = op32 @ X40                                  # one column per RHS, passes of at most 5
    op32.set_option("slot_cap", 0)
    assert rel(split, lanes) < 2e-6
    imps32 = [oracle.import_bf(T, T, op32.export(i)) for i in range(op32.n_factorizations())]
    want32 = np.stack([oracle.matvec_many(imps32, X40[:, j].astype(np.complex128), nthreads=8, rows=6000)
                       for j in (0, 17, 39)], axis=1)
    assert rel(split[:, [0, 17, 39]], want32) < 2e-6
    op32.set_option("slot_cap", 1)                      # not even one column: an error, not a fault
    with pytest.raises(bfgpu.ArgumentError):
        op32 @ X40
    op32.set_option("slot_cap", 0)


def test_matrix_core_rhs_matches_oracle(bfgpu, oracle, ctx):
    """Complex FP32 right-hand sides across lanes on the matrix cores
    (stacked 16-row tiles, items of more than 16 rows, partial sums and
    identity combines) against the oracle's FP64 product of the same FP32
    factors and against the vector-ALU kernel.  Bar: 2e-6 relative, the
    FP32 rounding of a few hundred products per entry along the level
    chain (the two FP32 kernels differ in summation order only)."""
    say("two squares N=4096 fp32: matrix-core right-hand sides")
    n = 4096
    obs, src = two_squares(n, seed=4)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    g = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(To, Ts), 0, 0, K, 1e-4,
                          acctype=np.complex64, ctx=ctx)
    # stored last-level blocks (the matrix-free last level has a test of its
    # own): the P level's items have 64 rows (a leaf), several row tiles each
    g.set_option("pfree", 0)
    imp = oracle.import_bf(To, Ts, g.export())
    rng = np.random.default_rng(12)
    X = np.asfortranarray(crandn(rng, n, 70)).astype(np.complex64)
    want = np.stack([imp.matvec(X[:, j].astype(np.complex128)) for j in range(70)], axis=1)
    assert g.get_option("mfma") == 32
    g.set_option("rhs_lanes", 16)
    for nr in (16, 17, 33, 64, 70):
        got = (g @ np.asfortranarray(X[:, :nr])).astype(np.complex128)
        assert rel(got, want[:, :nr]) < 2e-6, nr
    for rows in (32, 48, 64, 16):  # tiles of 1..4 row blocks
        g.set_option("mfma", rows)
        assert g.get_option("mfma") == rows
        got = (g @ np.asfortranarray(X[:, :33])).astype(np.complex128)
        assert rel(got, want[:, :33]) < 2e-6, rows
    Y0 = np.asfortranarray(crandn(rng, n, 17)).astype(np.complex64)
    a, b = 0.5 - 0.25j, -1.5 + 0.5j
    got = g.mul(Y0.copy(order="F"), np.asfortranarray(X[:, :17]), a, b).astype(np.complex128)
    assert rel(got, a * want[:, :17] + b * Y0.astype(np.complex128)) < 2e-6
    g.set_option("mfma", 0)
    valu = (g @ np.asfortranarray(X[:, :64])).astype(np.complex128)
    g.set_option("mfma", 1)
    assert rel(valu, want[:, :64]) < 2e-6
    # passes over column blocks: 70 = 16 + 16 + 16 + 16 + 6 across lanes,
    # and 20 = 7 + 7 + 6 with one column per RHS
    Yc = np.asfortranarray(crandn(rng, n, 70)).astype(np.complex64)
    g.set_option("rhs_chunk", 16)
    got = g.mul(Yc.copy(order="F"), np.asfortranarray(X), a, b).astype(np.complex128)
    g.set_option("rhs_chunk", 0)
    assert rel(got, a * want + b * Yc.astype(np.complex128)) < 2e-6
    g.set_option("rhs_lanes", 0)
    g.set_option("rhs_chunk", 7)
    got = (g @ np.asfortranarray(X[:, :20])).astype(np.complex128)
    g.set_option("rhs_chunk", 0)
    g.set_option("rhs_lanes", 16)
    assert rel(got, want[:, :20]) < 2e-6

    say("far field sphere N=6000 fp32: matrix-core right-hand sides, partial sums")
    pts = fibonacci_sphere(6000, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(T, T), K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                farfieldonly=True, ctx=ctx)
    assert op.launch_stats()[-1, 7] > 0
    op.set_option("pfree", 0)  # stored P: partial sums and their combine pass on the matrix cores
    imps = [oracle.import_bf(T, T, op.export(i)) for i in range(len(op.farinteractions))]
    Xs = np.asfortranarray(crandn(rng, 6000, 40)).astype(np.complex64)
    want = np.stack([oracle.matvec_many(imps, Xs[:, j].astype(np.complex128), nthreads=8, rows=6000)
                     for j in range(40)], axis=1)
    for nr in (32, 40):
        got = (op @ np.asfortranarray(Xs[:, :nr])).astype(np.complex128)
        assert rel(got, want[:, :nr]) < 2e-6, nr
    for rows in (64, 32, 16):
        op.set_option("mfma", rows)
        got = (op @ np.asfortranarray(Xs[:, :40])).astype(np.complex128)
        assert rel(got, want[:, :40]) < 2e-6, rows
    op.set_option("mfma", 0)
    valu = (op @ Xs).astype(np.complex128)
    assert rel(valu, want) < 2e-6


def test_matrix_free_last_level(bfgpu, oracle, ctx, tmp_path):
    """One-RHS FP32 products compute the last level's P entries (point-
    kernel entries between observer-leaf points and skeleton points) in the
    kernel, forward, adjoint and transpose: each product equals the stored-P
    product to FP32 rounding of the entries (1e-5), with alpha / beta; FP64
    operators and operators read from a file keep the stored blocks."""
    say("far field sphere N=6000 fp32: matrix-free last level")
    pts = fibonacci_sphere(6000, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(T, T)
    op = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4,
                                admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                farfieldonly=True, ctx=ctx)
    assert op.get_option("pfree") == 1
    rng = np.random.default_rng(11)
    x = crandn(rng, 6000).astype(np.complex64)
    y0 = crandn(rng, 6000).astype(np.complex64)
    on = op @ x
    on_ab = op.mul(y0.copy(), x, 0.5 - 1j, 2.0)
    import torch
    xd = torch.from_numpy(x).to("cuda:0")
    _, desc_on = op.launch_times(xd, torch.zeros_like(xd), iters=2)
    assert any(int(d[2]) == 6 for d in desc_on)
    # adjoint and transpose: the first transposed step writes the skeleton entries
    adj_on, tr_on = op.adjoint() @ x, op.transpose() @ x
    adj_ab = op.adjoint().mul(y0.copy(), x, 0.5 - 1j, 2.0)
    _, desc_adj = op.adjoint().launch_times(xd, torch.zeros_like(xd), iters=2)
    assert any(int(d[2]) == 7 for d in desc_adj)
    op.set_option("pfree", 0)
    assert op.get_option("pfree") == 0
    off = op @ x
    off_ab = op.mul(y0.copy(), x, 0.5 - 1j, 2.0)
    assert rel(on, off) < 1e-5
    assert rel(on_ab, off_ab) < 1e-5
    assert rel(adj_on, op.adjoint() @ x) < 1e-5
    assert rel(tr_on, op.transpose() @ x) < 1e-5
    assert rel(adj_ab, op.adjoint().mul(y0.copy(), x, 0.5 - 1j, 2.0)) < 1e-5
    # the point kernel is symmetric: transpose(F) x = F x up to the compression
    assert rel(tr_on, on) < 1e-3
    # the transposed copy covers the stored levels; the matrix-free level is
    # not copied and stays on its own kernel
    op.set_option("pfree", 1)
    op.set_option("tcopy", 1)
    assert op.get_option("tcopy") == 1
    with pytest.raises(bfgpu.ArgumentError):
        _force_tcopy(op, bfgpu.ButterflyFactorization.tile_shape(5, 0))
    _force_tcopy(op, bfgpu.ButterflyFactorization.tile_shape(5, 0), skip_last=True)
    assert rel(op.adjoint() @ x, adj_on) < 1e-5 and rel(op.transpose() @ x, tr_on) < 1e-5
    assert rel(op.adjoint().mul(y0.copy(), x, 0.5 - 1j, 2.0), adj_ab) < 1e-5
    op.set_option("tcopy", 0)
    op.set_option("pfree", 0)
    # the FP64 form of the same operator through the oracle's matvec
    imps = [oracle.import_bf(T, T, op.export(i)) for i in range(len(op.farinteractions))]
    ref = oracle.matvec_many(imps, x.astype(np.complex128), rows=6000)
    assert rel(on, ref) < 1e-5
    op.set_option("pfree", 1)
    # recompressed operators keep P (and the matrix-free level); files do not carry the points
    rc = op.recompress(1e-4)
    assert rc.get_option("pfree") == 1
    # the recompressed operator's computed-P products against its stored-P
    # products and against the oracle's product of its exported blocks
    rc_on, rc_adj_on = rc @ x, rc.adjoint() @ x
    rc.set_option("pfree", 0)
    rc_off, rc_adj_off = rc @ x, rc.adjoint() @ x
    rc.set_option("pfree", 1)
    assert rel(rc_on, rc_off) < 1e-5 and rel(rc_adj_on, rc_adj_off) < 1e-5
    rimps = [oracle.import_bf(T, T, rc.export(i)) for i in range(rc.n_factorizations())]
    x64 = x.astype(np.complex128)
    assert rel(rc_on, oracle.matvec_many(rimps, x64, rows=6000)) < 1e-5
    radj = [b.transform(1) for b in rimps]
    assert rel(rc_adj_on, oracle.matvec_many(radj, x64, rows=6000)) < 1e-5
    path = str(tmp_path / "s.bfgpu")
    op.save(path)
    assert bfgpu.load(path, bt, ctx=ctx).get_option("pfree") == 0
    op64 = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bt, K, tol=1e-4,
                                  admissibility=bfgpu.CenterDistanceAdmissibility(2.0), farfieldonly=True, ctx=ctx)
    assert op64.get_option("pfree") == 0


def test_matrix_free_last_level_many_rhs(bfgpu, oracle, ctx):
    """Right-hand sides across lanes (16-64 per pass) compute the last
    level's point-kernel entries on the fly and multiply them on the matrix
    cores (one wavefront per observer leaf, up to four 16-row blocks): the
    products equal the stored-P matrix-core products to FP32 rounding of the
    entries (1e-5, the one-RHS matrix-free bar) and the oracle's FP64
    product of the same factors (1e-5); alpha / beta, two passes (70 RHS),
    leaves of a few to 64 rows (the sphere) and of more than 64 rows (the
    two squares' balanced trees, taken 64 rows at a time)."""
    import torch
    rng = np.random.default_rng(31)
    say("far field sphere N=6000 fp32 and two squares N=4096 fp32: matrix-free last level, many RHS")
    pts = fibonacci_sphere(6000, 4.0)
    T = bfgpu.ClusterTree(pts, dims=3, leaf_size=64, mode="maxpoints")
    sphere = bfgpu.PetrovGalerkinBF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(T, T), K, tol=1e-4,
                                    admissibility=bfgpu.CenterDistanceAdmissibility(2.0), acctype=np.complex64,
                                    farfieldonly=True, ctx=ctx)
    obs, src = two_squares(4096, seed=5)
    To = bfgpu.ClusterTree(obs, dims=2, leaf_size=64)
    Ts = bfgpu.ClusterTree(src, dims=2, leaf_size=64)
    sq = bfgpu.assemble_BF(bfgpu.HelmholtzKernelMatrix(K), bfgpu.BlockTree(To, Ts), 0, 0, K, 1e-4,
                           acctype=np.complex64, ctx=ctx)
    for name, op, n, trees in (("sphere", sphere, 6000, (T, T)), ("squares", sq, 4096, (To, Ts))):
        assert op.get_option("pfree") == 1 and op.get_option("mfma_emul") == 1
        op.set_option("mfma_emul", 0)  # FP32 matrix cores first
        X = np.asfortranarray(crandn(rng, n, 70)).astype(np.complex64)
        Y0 = np.asfortranarray(crandn(rng, n, 33)).astype(np.complex64)
        a, b = 0.5 - 0.25j, -1.5 + 0.5j
        on = op @ X
        on_ab = op.mul(Y0.copy(order="F"), np.asfortranarray(X[:, :33]), a, b)
        Xd = torch.from_numpy(np.ascontiguousarray(X[:, :40].T)).to("cuda:0").t()
        _, desc = op.launch_times(Xd, torch.zeros_like(Xd), iters=1, nrhs=40)
        assert any(int(d[2]) == 10 for d in desc), name
        op.set_option("pfree", 0)
        off = op @ X
        off_ab = op.mul(Y0.copy(order="F"), np.asfortranarray(X[:, :33]), a, b)
        op.set_option("pfree", 1)
        say(name, "matrix-free vs stored P", rel(on, off), "alpha/beta", rel(on_ab, off_ab))
        assert rel(on, off) < 1e-5 and rel(on_ab, off_ab) < 1e-5, name
        nb = len(op.farinteractions) if hasattr(op, "farinteractions") else 1
        imps = [oracle.import_bf(trees[0], trees[1], op.export(i)) for i in range(nb)]
        cols = (0, 37, 69)
        want = np.stack([oracle.matvec_many(imps, X[:, j].astype(np.complex128), nthreads=8, rows=n)
                         for j in cols], axis=1)
        assert rel(on[:, cols].astype(np.complex128), want) < 1e-5, name
        # FP32 emulated on the bf16 matrix cores (bf16x3): operands carry
        # about 17 bits (|x - x_h - x_l| <= 2^-16 |x|), so the bar is 1e-4,
        # the compression tolerance, and the measured difference is printed
        op.set_option("mfma_emul", 1)
        assert op.get_option("mfma_emul") == 1
        em = op @ X
        em_ab = op.mul(Y0.copy(order="F"), np.asfortranarray(X[:, :33]), a, b)
        op.set_option("mfma_emul", 0)
        say(name, "bf16x3 vs FP32 matrix cores", rel(em, on), "vs oracle", rel(em[:, cols].astype(np.complex128), want))
        assert rel(em, on) < 1e-4 and rel(em_ab, on_ab) < 1e-4, name
        assert rel(em[:, cols].astype(np.complex128), want) < 1e-4, name
