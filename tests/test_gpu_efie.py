"""GPU parity of the EFIE / RWG kernel plug-in against the oracle
(oracle/efie.py): the device fill to 1e-13 (FP64) and 2e-6 (FP32); a
far-field operator factorized with it applies as the oracle's matvec on the
same blocks (FP64, 1e-12) and matches dense far-field rows of the oracle's
EFIE entries to ten times the tolerance; the near-field quadrature
(singularity extraction) equals the oracle's entries to 1e-12, and the
whole operator (near + far) applies as the dense matrix of its near blocks
and far entries to ten times the tolerance."""
import sys

import numpy as np
import pytest

from conftest import crandn

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


def test_efie_fill_matches_oracle(bfgpu, ctx):
    from bfgpu import rwg
    from oracle import efie
    V, T = rwg.icosphere(1.0, 3)
    km = rwg.EFIEKernelMatrix(rwg.RWG(V, T), K)
    n = len(km.rwg)
    rng = np.random.default_rng(0)
    say("EFIE fill, icosphere(3),", n, "functions")
    for m_, n_ in ((1, 1), (7, 13), (64, 40)):
        rows = rng.choice(n, m_, replace=False).astype(np.int32)
        cols = rng.choice(n, n_, replace=False).astype(np.int32)
        want = efie.efie_block(V, T, K, rows, cols)
        got = km.block(rows, cols, ctx=ctx)
        assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want), (m_, n_)
        got32 = km.block(rows, cols, acctype=np.complex64, ctx=ctx).astype(np.complex128)
        assert np.linalg.norm(got32 - want) <= 2e-6 * np.linalg.norm(want), (m_, n_)
    # other rules and weights reach the device unchanged
    km2 = rwg.EFIEKernelMatrix(km.rwg, K, alpha=0.3 + 0.1j, beta=-2.0j, rules=(4, 1))
    rows, cols = np.arange(0, n, 97, dtype=np.int32), np.arange(3, n, 71, dtype=np.int32)
    want = efie.efie_block(V, T, K, rows, cols, rules=(4, 1), alpha=0.3 + 0.1j, beta=-2.0j)
    assert rel(km2.block(rows, cols, ctx=ctx), want) < 1e-13
    with pytest.raises(bfgpu.ArgumentError):
        km.block(np.array([n], np.int32), cols, ctx=ctx)


def _far_rows(op, V, T, x, rows):
    from oracle import efie
    Tr = op.tree.testcluster
    S = op.tree.trialcluster
    by_obs = {}
    for o, s in op.farinteractions:
        by_obs.setdefault(int(o), []).append(int(s))
    leaf_of = {}
    for v in Tr.leaves():
        for i in Tr.cluster_values(v):
            leaf_of[int(i)] = v
    out = np.zeros(len(rows), complex)
    for j, r in enumerate(rows):
        node = leaf_of[int(r)]
        while node >= 0:
            for s in by_obs.get(node, ()):
                cols = S.cluster_values(s)
                out[j] += efie.efie_block(V, T, K, [r], cols)[0] @ x[cols]
            node = Tr.cluster_parent(node)
    return out


def test_efie_far_field_operator(bfgpu, oracle, ctx):
    from bfgpu import rwg
    V, T = rwg.icosphere(1.5, 4)
    fns = rwg.RWG(V, T)
    n = len(fns)
    Tree = bfgpu.ClusterTree(fns.centers, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(Tree, Tree)
    km = rwg.EFIEKernelMatrix(fns, K)
    say("EFIE far field, icosphere(4) radius 1.5,", n, "functions, fp64 tol 1e-4")
    op = bfgpu.PetrovGalerkinBF(km, bt, K, tol=1e-4, admissibility=bfgpu.CenterDistanceAdmissibility(2.0),
                                farfieldonly=True, ctx=ctx)
    assert op.farinteractions.shape[0] > 50
    rng = np.random.default_rng(1)
    x = crandn(rng, n)
    y = op @ x
    # the GPU's own blocks through the oracle's matvec
    imps = [oracle.import_bf(Tree, Tree, op.export(i)) for i in range(len(op.farinteractions))]
    assert rel(y, oracle.matvec_many(imps, x, rows=n)) < 1e-12
    # accuracy: dense far-field rows of the oracle's entries
    rows = np.sort(rng.choice(n, 12, replace=False))
    ye = _far_rows(op, V, T, x, rows)
    assert rel(y[rows], ye) < 10 * 1e-4
    # FP32 factorization of the same operator
    op32 = bfgpu.PetrovGalerkinBF(km, bt, K, tol=1e-4, admissibility=bfgpu.CenterDistanceAdmissibility(2.0),
                                  acctype=np.complex64, farfieldonly=True, ctx=ctx)
    y32 = op32 @ x.astype(np.complex64)
    assert rel(y32[rows].astype(complex), ye) < 10 * 1e-4
    assert rel(op.adjoint() @ x, op.adjoint().materialize() @ x) < 1e-12


def test_efie_near_quadrature_matches_oracle(bfgpu, ctx):
    from bfgpu import rwg
    from oracle import efie
    V, T = rwg.icosphere(1.0, 3)
    fns = rwg.RWG(V, T)
    c = fns.centers
    km = rwg.EFIEKernelMatrix(fns, K)
    say("EFIE near quadrature, icosphere(3),", len(fns), "functions")
    for seed_fn in (0, 777):
        near = np.argsort(np.linalg.norm(c - c[seed_fn], axis=1))[:12].astype(np.int32)
        want = efie.efie_near_block(V, T, K, near, near)
        got = km.near_block(near, near, ctx=ctx)
        assert rel(got, want) < 1e-12, seed_fn
        got32 = km.near_block(near, near, acctype=np.complex64, ctx=ctx).astype(np.complex128)
        assert rel(got32, want) < 2e-6
    # other rule degrees and extraction distance reach the device unchanged
    km2 = rwg.EFIEKernelMatrix(fns, K, near_rules=(3, 4), near_far_rules=(2, 2), near_factor=1.5)
    far = np.argsort(np.linalg.norm(c - c[5], axis=1))[:20].astype(np.int32)
    want = efie.efie_near_block(V, T, K, far[:9], far, near_rules=(3, 4), rules=(2, 2), near_factor=1.5)
    assert rel(km2.near_block(far[:9], far, ctx=ctx), want) < 1e-12


def test_efie_whole_operator_with_near_field(bfgpu, ctx):
    from bfgpu import rwg
    from oracle import efie
    V, T = rwg.icosphere(1.0, 3)
    fns = rwg.RWG(V, T)
    n = len(fns)
    Tree = bfgpu.ClusterTree(fns.centers, dims=3, leaf_size=64, mode="maxpoints")
    bt = bfgpu.BlockTree(Tree, Tree)
    km = rwg.EFIEKernelMatrix(fns, K)
    say("EFIE whole operator (near + far), icosphere(3),", n, "functions, fp64 tol 1e-4")
    op = bfgpu.PetrovGalerkinBF(km, bt, K, tol=1e-4, admissibility=bfgpu.CenterDistanceAdmissibility(2.0), ctx=ctx)
    N = op.nearinteractions
    assert N is not None and op.farinteractions.shape[0] > 0
    # near blocks: the oracle's near quadrature, on the smallest blocks
    pairs = op.nearpairs
    sizes = [len(Tree.cluster_values(o)) * len(Tree.cluster_values(s_)) for o, s_ in pairs]
    for i in np.argsort(sizes)[:3]:
        o, s_ = pairs[i]
        want = efie.efie_near_block(V, T, K, Tree.cluster_values(o), Tree.cluster_values(s_))
        assert rel(N.block(int(i)), want) < 1e-12
    # dense operator: near blocks where the pairs are near, the far
    # quadrature's entries where they are far
    A = np.zeros((n, n), complex)
    covered = np.zeros((n, n), np.int32)
    for i, (o, s_) in enumerate(pairs):
        r, c_ = Tree.cluster_values(o), Tree.cluster_values(s_)
        A[np.ix_(r, c_)] = N.block(i)
        covered[np.ix_(r, c_)] += 1
    Zfar = km.block(np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32), ctx=ctx)
    for o, s_ in op.farinteractions:
        r, c_ = Tree.cluster_values(o), Tree.cluster_values(s_)
        A[np.ix_(r, c_)] = Zfar[np.ix_(r, c_)]
        covered[np.ix_(r, c_)] += 1
    assert np.all(covered == 1)  # near and far tile the matrix once
    rng = np.random.default_rng(3)
    x = crandn(rng, n)
    Nd = np.zeros((n, n), complex)
    for i, (o, s_) in enumerate(pairs):
        Nd[np.ix_(Tree.cluster_values(o), Tree.cluster_values(s_))] = N.block(i)
    assert rel(N @ x, Nd @ x) < 1e-12
    assert rel(N.adjoint() @ x, Nd.conj().T @ x) < 1e-12
    assert rel(N.transpose() @ x, Nd.T @ x) < 1e-12
    e = rel(op @ x, A @ x)
    say("whole operator vs dense", e)
    assert e < 10 * 1e-4
    assert rel(op.adjoint() @ x, A.conj().T @ x) < 10 * 1e-4
    op32 = bfgpu.PetrovGalerkinBF(km, bt, K, tol=1e-4, admissibility=bfgpu.CenterDistanceAdmissibility(2.0),
                                  acctype=np.complex64, ctx=ctx)
    e32 = rel((op32 @ x.astype(np.complex64)).astype(complex), A @ x)
    say("fp32 whole operator vs dense", e32)
    assert e32 < 10 * 1e-4
